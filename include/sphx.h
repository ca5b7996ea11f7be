/*
 * sphx.h - C ABI of the MI355X-native particle hot path (spatial-hash grid,
 * SPH/IISPH and multiphase neighbour loops, MPM transfer, RPY mobility).
 *
 * Two layers are exported from libsphx.so:
 *
 *   1. The legacy drop-in entry points, with the names, argument meaning and
 *      call order of the reference's host driver
 *      (SPH_Solver/fluids/fluid_system_host.cuh:44-107).  They act on one
 *      process-wide default context, synchronise after every call and latch
 *      errors (printed to stderr, readable with sphx_last_error()), exactly
 *      like the reference's void wrappers, which print cudaGetLastError()
 *      and carry on.  A C++ reference `T&` argument is passed as `T*` at the
 *      ABI level, so the reference's own declarations link against these.
 *
 *   2. The context API (sphx_*): explicit context objects, int status
 *      returns (0 = success, <0 = error with a thread-local message),
 *      asynchronous work on the context's HIP stream.
 *
 * All particle data crosses the boundary in the reference's AoS formats
 * (displayPack 32 B, calculationPack 224 B, ParamCarrier 464 B); the device
 * layout behind them is private (structure of arrays, cell-sorted).
 */
#ifndef SPHX_H
#define SPHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Types (layout-identical to SPH_Solver/catpaw/geometry.h:48-176 and        */
/* SPH_Solver/common_header.h:28-137; every field is 4 bytes, no padding).   */
/* ------------------------------------------------------------------------ */

typedef struct { float x, y, z; } sphx_float3;
typedef struct { float x, y, z, w; } sphx_float4;
typedef struct { int x, y, z; } sphx_int3;
typedef struct { float data[9]; } sphx_mat3; /* row-major, geometry.h:142-176 */

#define SPHX_MAX_FLUIDNUM 3   /* common_header.h:17 MAX_FLUIDNUM / FLUID_NUM */
#define SPHX_TYPE_FLUID 0     /* common_header.h:20-23 */
#define SPHX_TYPE_BOUNDARY 1
#define SPHX_TYPE_ELASTIC 2
#define SPHX_TYPE_GRANULAR 3
#define SPHX_GRID_UNDEF 0xFFFFFFFFu /* fluid_system_kern.cuh:16 GRID_UNDEF */

/* common_header.h:28-109 ParamCarrier */
typedef struct sphx_ParamCarrier {
  float acclimit;
  float vlimit;
  sphx_float3 gravity;
  sphx_float3 volmin;
  sphx_float3 volmax;
  sphx_float3 softminx;
  sphx_float3 softmaxx;
  int num;
  int maxNum;
  float dt;
  float simscale;
  float viscosity;
  float restdensity;
  float mass;
  float radius;
  float smoothradius;
  float scalep;
  float massArr[SPHX_MAX_FLUIDNUM];
  float densArr[SPHX_MAX_FLUIDNUM];
  float viscArr[SPHX_MAX_FLUIDNUM];
  float boundstiff;
  float kpoly6;
  float kspiky;
  float kspikydiff;
  float klaplacian;
  float kspline;
  float bmass;
  float bRestdensity;
  float bvisc;
  float cellsize;
  int searchnum;
  int neighbornum;
  int neighborid[27];
  sphx_float3 gridmin, gridmax;
  sphx_float3 gridsize;
  sphx_float3 gridIdfac;
  sphx_int3 gridres;
  int gridtotal;
  float mpmcellsize;
  sphx_float3 mpmXmin, mpmXmax;
  sphx_int3 mpmRes;
  int mpmNodeNum;
  float solidK;
  float solidG;
  float a_phi;
  float a_psi;
  float k_c;
  float w_deltax;
  float fijn;
  float SFlipControl;
  float LFlipControl;
  int mpmSplitNum;
  float collisionStiff;
  float surfaceTensionK;
  float surfaceTensionFluidC;
  float fBlending;
} sphx_ParamCarrier;

/* common_header.h:112-116 displayPack */
typedef struct sphx_displayPack {
  sphx_float3 pos;
  sphx_float4 color;
  int type;
} sphx_displayPack;

/* common_header.h:118-137 calculationPack.  `dens` holds 1/rho after the
 * WCSPH density pass, rho after the IISPH density pass, and the boundary
 * volume for boundary particles (fluid_system_kern.cu:276,306,1114). */
typedef struct sphx_calculationPack {
  sphx_float3 vel;
  sphx_float3 veleval;
  float pressure;
  float dens;
  float restdens;
  float mass;
  float visc;
  int bornid;
  sphx_float3 accel;
  sphx_mat3 deformGrad;
  sphx_mat3 stress;
  sphx_mat3 B;
  float solidG;
  float solidK;
  sphx_float3 X;
  sphx_mat3 invA;
} sphx_calculationPack;

/* SPH_Solver/stokesian/Mobility.cuh:6-16 stokesianBufList (device pointers). */
typedef struct sphx_stokesianBufList {
  sphx_float3* cuUnew;
  sphx_displayPack* dispBuffer;
  sphx_float3* cuOmega;
  sphx_float3* cuForce;
  float dt;
  float simscale;
} sphx_stokesianBufList;

/* Solver choices that the reference fixes in code.  The defaults reproduce
 * the reference; every other value is a documented extension.  No
 * counterpart in the reference (it hard-codes each of these). */
typedef struct sphx_SolverOptions {
  /* IISPH (fluid_system_kern.cu:730-1050, fluid_system_host.cu:525-552) */
  float omega;              /* relaxed-Jacobi weight; reference 0.1 (:1026) */
  float warm_start;         /* p0 = warm_start * p(t-1); reference 0.5 (:899) */
  int full_support;         /* 0: pressure loops test r < h as the reference
                               (:750,:822,:867,:931,:985); 1: r < 2h */
  int boundary_pressure;    /* 1: enable the boundary terms the reference has
                               commented out (:763-764, :833-834, :1000-1002)
                               plus the matching a_ii term; 0: reference */
  float tol;                /* stop when mean |rho-rho0|/rho0 <= tol (FP64);
                               the reference loop never stops (A1) */
  int min_iters;            /* reference criterion needs iter >= 2 (:545) */
  int max_iters;
  int clamp_to_grid;        /* 1: integration keeps particles inside the
                               valid cell range (guard for long runs) */
  /* WCSPH kernel family / EOS (fluid_system.cpp:308-313, kern.cu:310) */
  int wcsph_kernel;         /* 0: poly6 + spiky, support h (reference);
                               1: cubic spline, support 2h */
  float eos_B;              /* Tait stiffness B; <= 0 selects 2.5*rho0 */
  float eos_gamma;          /* Tait exponent; reference 7 */
  float art_visc_alpha;     /* >0: Monaghan artificial viscosity between
                               fluid pairs with sound speed c0 */
  float sound_speed;        /* c0 for art_visc_alpha */
  /* multiphase (multiphase.cu spec) */
  float drift_relax;        /* tau numerator: tau = drift_relax / nu_i */
  float drift_diffusion;    /* sigma_D of the alpha-gradient term */
  float drift_pressure_scale; /* cont*cont2 of multiphase.cu:133 */
  float drift_diffusion_scale;/* cont*cont1 of multiphase.cu:162 */
  int boussinesq;           /* 1: mixture buoyancy (rho_m-rho_ref)/rho_ref*g */
  float rho_ref;            /* reference density of the mixture solve */
  int clamp_pressure;       /* 1: p = max(p, 0) after each iteration (the
                               clamp the reference has commented out at
                               kern.cu:1035-1036); the density error is then
                               the complementarity residual: |rho-rho0|/rho0
                               where the iterate's p > 0, max(rho-rho0,0)/rho0
                               where p = 0.  0: reference, |rho-rho0|/rho0 */
  int reserved[7];
} sphx_SolverOptions;

/* ------------------------------------------------------------------------ */
/* Errors                                                                    */
/* ------------------------------------------------------------------------ */

/* Message of the last failing call on this thread ("" if none). */
const char* sphx_last_error(void);
/* Number of HIP devices visible (0 without a GPU; never fails). */
int sphx_device_count(void);
/* Device buffers for callers of the entry points that take device pointers
 * (getMobU, getMobU_walkthrough, sphx_rpy_apply_device); kind: 1 host to
 * device, 2 device to host, 3 device to device. */
int sphx_dev_alloc(void** out, size_t bytes);
int sphx_dev_free(void* p);
int sphx_dev_copy(void* dst, const void* src, size_t bytes, int kind);
/* Library version string. */
const char* sphx_version(void);

/* ------------------------------------------------------------------------ */
/* Legacy drop-in entry points (fluid_system_host.cuh:44-107)                */
/* ------------------------------------------------------------------------ */

void cudaInit(int argc, char** argv);                 /* host.cu:30 */
void cudaExit(int argc, char** argv);                 /* host.cu:26 */
void FluidClearCUDA(void);                            /* host.cu:79 */
void FluidSetupCUDA(sphx_ParamCarrier* params);       /* host.cu:103 */
void CarryParam(sphx_ParamCarrier* hostCarrier);      /* kern.cu:23 */
void computeNumBlocks(int numPnts, int maxThreads, int* numBlocks,
                      int* numThreads);               /* host.cu:70 */
void GetParticleIndexCUDA(void);                      /* host.cu:147 */
void GetGridListCUDA(void);                           /* host.cu:161 */
void RearrageDataCUDA(void);                          /* host.cu:187 [sic] */
void ComputeBoundaryDensity(void);                    /* host.cu:510 */
void MfComputePressureCUDA(void);                     /* host.cu:337 */
void MfComputeForceCUDA(void);                        /* host.cu:421 */
void MfAdvanceCUDA(void);                             /* host.cu:385 */
void ComputeDensityIISPH_CUDA(void);                  /* host.cu:505 */
void PredictAdvection(void);                          /* host.cu:515 */
void PressureSolve(void);                             /* host.cu:525 */
void Integration(void);                               /* host.cu:554 */
/* SPH elastic solid (fluid_system_host.cuh:96-100).  The reference's
 * hyperelastic _X kernels (sph_solid.cu:650-661) are empty, so the two X
 * entry points do nothing, as there. */
void ComputeSolidTensorCUDA(void);                    /* host.cu:469 */
void ComputeSolidForceCUDA(void);                     /* host.cu:475 */
void ComputeSolidTensorX_CUDA(void);                  /* host.cu:481 */
void ComputeSolidForceX_CUDA(void);                   /* host.cu:487 */
/* host.cu:492-501: grid hash, sort and reorder, then ComputeInvA, whose
 * body does nothing past its guards (sph_solid.cu:662-675). */
void InitializeSolid_CUDA(void);

/* Host-side transfers the reference performs with raw cudaMemcpy on the
 * global buffers (fluid_system.cpp:979-999). */
void sphx_legacy_upload(const sphx_displayPack* disp,
                        const sphx_calculationPack* calc, int start, int end);
void sphx_legacy_download(sphx_displayPack* disp, sphx_calculationPack* calc,
                          int count);
/* The default context behind the legacy entry points (NULL before
 * FluidSetupCUDA). */
struct sphx_context* sphx_legacy_context(void);

/* ------------------------------------------------------------------------ */
/* Context API                                                               */
/* ------------------------------------------------------------------------ */

typedef struct sphx_context sphx_context;

/* Allocate device buffers for p->maxNum particles and the grid p->gridtotal
 * (fluid_system_host.cu:103-144); copies the parameter block (CarryParam). */
int sphx_create(sphx_context** out, const sphx_ParamCarrier* p);
int sphx_destroy(sphx_context* ctx);
/* CarryParam: replace the parameter block (num, dt, grid ...). Grid size
 * changes reallocate the cell table. */
int sphx_set_params(sphx_context* ctx, const sphx_ParamCarrier* p);
int sphx_get_params(const sphx_context* ctx, sphx_ParamCarrier* p);
void sphx_default_options(sphx_SolverOptions* o);
int sphx_set_options(sphx_context* ctx, const sphx_SolverOptions* o);
/* Block until every queued kernel of the context has finished. */
int sphx_sync(sphx_context* ctx);

/* Particle transfers (AoS at the boundary).  upload writes slots
 * [start, start+count) and sets num = max(num, start+count) when grow != 0. */
int sphx_upload(sphx_context* ctx, const sphx_displayPack* disp,
                const sphx_calculationPack* calc, int start, int count,
                int grow);
int sphx_download(sphx_context* ctx, sphx_displayPack* disp,
                  sphx_calculationPack* calc, int start, int count);
int sphx_num_particles(const sphx_context* ctx);

/* Grid build (fluid_system_kern.cu:36-121).  Each pointer may be NULL.
 * mgcell/midsort/mgndx have num entries; gridoff/gridcnt gridtotal entries;
 * MFidTable maps bornid -> sorted slot (num entries, -1 when absent). */
int sphx_download_grid(sphx_context* ctx, uint32_t* mgcell, uint32_t* midsort,
                       uint32_t* mgndx, int32_t* gridoff, int32_t* gridcnt,
                       int32_t* MFidTable);
int sphx_grid_hash(sphx_context* ctx);      /* InitialSort   kern.cu:36 */
int sphx_grid_sort(sphx_context* ctx);      /* sort_by_key + CalcFirstCnt +
                                               GetCnt host.cu:161-185 */
int sphx_grid_reorder(sphx_context* ctx);   /* RearrangeData kern.cu:98 */
int sphx_grid_build(sphx_context* ctx);     /* the three above */

/* WCSPH (fluid_system_kern.cu:249-534) */
int sphx_boundary_volume(sphx_context* ctx);        /* :249 */
int sphx_wcsph_density_pressure(sphx_context* ctx); /* :279 */
int sphx_wcsph_force(sphx_context* ctx);            /* :397 */
int sphx_wcsph_advance(sphx_context* ctx);          /* :426 */
/* sort + density/pressure + force + advance (fluid_system.cpp:514-539) */
int sphx_wcsph_step(sphx_context* ctx);

/* SPH elastic solid (sph_solid.cu) and surface tension, after the WCSPH
 * density pass (dens = 1/rho).  Deformation gradients and stresses are the
 * calculationPack's deformGrad and stress. */
int sphx_solid_tensor(sphx_context* ctx);           /* ComputeSolidTensor :521 */
int sphx_solid_force(sphx_context* ctx);            /* ComputeSolidForce :612 */
/* SurfaceTension (fluid_system_kern.cu:605; its launch is commented out in
 * MfComputeForceCUDA, host.cu:433): accel += sum_j C cos(1.5 pi r/tR)/r x_ij
 * over 0 < |x_ij|^2 <= tR, tR = surfaceTensionK * smoothradius. */
int sphx_surface_tension(sphx_context* ctx);

/* IISPH (fluid_system_kern.cu:635-1115, fluid_system_host.cu:505-556) */
int sphx_iisph_density(sphx_context* ctx);          /* :1093 */
int sphx_iisph_predict_advection(sphx_context* ctx);/* :768 + :883 */
/* One Pressure_DP + Pressure_Iter + press_l <- press_l1 (:944, :1009,
 * host.cu:530-536).  Optionally returns the mean signed residual over all
 * num slots (host.cu:539-541, kg/m^3) and the FP64 mean |rho-rho0|/rho0
 * over fluid particles. Either pointer may be NULL (then no host sync). */
int sphx_iisph_pressure_iteration(sphx_context* ctx, double* mean_signed,
                                  double* mean_abs_rel);
/* Iterate to options.tol (min_iters, max_iters); returns iterations done and
 * the final FP64 error.  The convergence test runs on the device. */
int sphx_iisph_pressure_solve(sphx_context* ctx, int* iters, double* err);
int sphx_iisph_integrate(sphx_context* ctx);        /* :1052 */
/* sort + density + predict + solve + integrate (fluid_system.cpp:567-588
 * with Integration enabled). */
int sphx_iisph_step(sphx_context* ctx, int* iters, double* err);
/* The two kernels of sphx_iisph_predict_advection, callable apart so a
 * caller can add accelerations to the advected velocity in between. */
int sphx_iisph_predict_dii(sphx_context* ctx);      /* ComputeDii :768 */
int sphx_iisph_predict_aii(sphx_context* ctx);      /* ComputeAii :883 */
/* ComputeDensityIISPH (:1093) then ComputeDii (:768) in one neighbour
 * pass: density, advected velocity and d_ii, as the two calls back to back
 * (the composite steps use it; calls that change x, v.xyz, m or rho0 may
 * not sit between the two). */
int sphx_iisph_density_dii(sphx_context* ctx);
/* Diagnostics of the neighbour list the loops above share (no reference
 * counterpart: the reference scans the 27 cells in every loop).  counts[i]
 * = neighbours j != i within 2h of particle i (sorted order), -1 where the
 * list overflowed and the loops scan the cells instead; *cap = capacity.
 * Boundary and out-of-grid particles report 0. */
int sphx_nbr_counts(sphx_context* ctx, int* counts, int* cap);
/* IISPH per-slot buffers (any pointer may be NULL); vectors are 3*num. */
int sphx_download_iisph(sphx_context* ctx, float* dii, float* dijpj,
                        float* aii, float* rho_adv, float* press_l,
                        float* press_l1, float* residue);
int sphx_upload_iisph(sphx_context* ctx, const float* dii, const float* dijpj,
                      const float* aii, const float* rho_adv,
                      const float* press_l);

/* ------------------------------------------------------------------------ */
/* Multiphase drift-flux mixture (formula spec: SPH_Solver/fluids/           */
/* multiphase.cu, not compiled in the reference).                            */
/* ------------------------------------------------------------------------ */

/* Enable nphase (1..3) phases with rest densities rho[k] and kinematic
 * viscosities nu[k] (ParamCarrier densArr/viscArr semantics). */
int sphx_mf_setup(sphx_context* ctx, int nphase, const float* rho, const float* nu);
/* Volume fractions [count][nphase] for slots [start, start+count); sets the
 * mixture density/viscosity and g - a = g. */
int sphx_mf_upload(sphx_context* ctx, const float* alpha, int start, int count);
/* alpha [num][nphase], rho_m [num], drift [num][nphase][3], mforce [num][3];
 * any may be NULL. */
int sphx_mf_download(sphx_context* ctx, float* alpha, float* rho_m, float* drift,
                     float* mforce);
int sphx_mf_drift_velocity(sphx_context* ctx);   /* mfComputeDriftVel :78 */
int sphx_mf_alpha_advance(sphx_context* ctx);    /* mfComputeAlphaAdvance :218 */
int sphx_mf_correction(sphx_context* ctx);       /* mfComputeCorrection :275 */
/* drift stress (contributeForce_projectu :555-562) + Boussinesq buoyancy
 * added to the advected velocity (after sphx_iisph_predict_dii). */
int sphx_mf_force(sphx_context* ctx);
/* Full two-phase step (sort, density, drift, alpha, correction, predict,
 * mixture force, a_ii, pressure solve, integrate). */
int sphx_mf_step(sphx_context* ctx, int* iters, double* err);

/* Legacy multiphase wrappers: declared by the reference
 * (fluid_system_host.cuh:62-64,77) with bodies commented out
 * (fluid_system_host.cu:353-419); here they run the stages above. */
void MfComputeDriftVelCUDA(void);
void MfComputeAlphaAdvanceCUDA(void);
void MfComputeCorrectionCUDA(void);
void ComputeForceCUDA_ProjectU(void);

/* ------------------------------------------------------------------------ */
/* MPM transfer with Drucker-Prager plasticity                               */
/* (SPH_Solver/fluids/MpmSolver.cu; node grid of ParamCarrier mpmXmin,       */
/* mpmcellsize, mpmRes, node id GetMpmIdx :35-38 = (y*resz + z)*resx + x).   */
/* ------------------------------------------------------------------------ */

/* Allocate the node grid of the current params (MpmAllocateBufferCUDA
 * host.cu:228).  kernel: 0 the reference's cubic B-spline (:155-195), 1 a
 * quadratic B-spline.  physical: 0 reference gradients in cell units and
 * node force sigma grad N / rho (:700); 1 gradients in metres and node force
 * m_p / (rho_p m_I) sigma grad N.  Fails if an SPH grid cell is narrower
 * than the spline support (the reference's node sums would miss
 * particles). */
int sphx_mpm_setup(sphx_context* ctx, int kernel, int physical);
/* MpmParticleToGrid_APIC (apic=1, :823) / MpmParticleToGrid (apic=0, :197) */
int sphx_mpm_p2g(sphx_context* ctx, int apic);
int sphx_mpm_stress(sphx_context* ctx);          /* MpmParticleStress :477 */
int sphx_mpm_node_update(sphx_context* ctx);     /* MpmNodeUpdate :658 */
/* MpmParticleUpdate_APIC (apic=1, :891) / MpmParticleUpdate (apic=0, :706) */
int sphx_mpm_g2p(sphx_context* ctx, int apic);
/* RunMPM (fluid_system.cpp:541-564): sort, P2G, stress, node update, G2P */
int sphx_mpm_substep(sphx_context* ctx);
/* Node velocities [N][3] and masses [N] (either may be NULL). */
int sphx_mpm_download_nodes(sphx_context* ctx, float* v, float* mass);
/* Set node velocities and masses (marks every node block active). */
int sphx_mpm_upload_nodes(sphx_context* ctx, const float* v, const float* mass);
/* Number of 4^3-node blocks the last transfer touched. */
int sphx_mpm_active_blocks(sphx_context* ctx, int* count);
/* MpmGetMomentum (MpmSolver.cu:107-153): every node's velocity becomes the
 * trilinear-weighted mean velocity of the particles within one node
 * spacing on each axis, searched in the 27 SPH cells around the node's
 * cell (initMpm :39-73); 0 where the weights sum to <= 1e-6 or the node's
 * cell is outside the valid cells.  Node masses are kept.  Needs a built
 * grid. */
int sphx_mpm_get_momentum(sphx_context* ctx);
/* MpmColorTest (MpmSolver.cu:75-105): particles in the 27 cells around a
 * node left of x = 0 get colour (1, 1, 1, 1).  Needs a built grid. */
int sphx_mpm_color_test(sphx_context* ctx);

/* ---------------------------------------------------------------------- */
/* FP64 conservation totals (no counterpart in the reference; the build's
 * global mass / momentum reductions).  Per scene (1, or every scene of a
 * batched ensemble), over the in-grid fluid particles:
 *   [0] sum m, [1..3] sum m v, [4] sum m |v|^2 / 2, [5..7] sum m x
 * (centre of mass = [5..7] / [0]).  out holds max(1, nscenes) * 8 doubles.
 * A single scene's totals are reproducible bit for bit. */
#define SPHX_TOTALS_Q 8
int sphx_conservation_totals(sphx_context* ctx, double* out);
/* on: every integration (eager or captured) appends its totals to a device
 * ring of the last 4096 steps; the ring restarts at 0. */
int sphx_conservation_track(sphx_context* ctx, int on);
/* The last min(recorded, max_steps) steps of the ring, oldest first, as
 * [steps][scenes][8]; *steps = how many were written. */
int sphx_conservation_history(sphx_context* ctx, double* out, int max_steps, int* steps);

/* Legacy MPM entry points (fluid_system_host.cuh; host.cu:228-296). */
void MpmAllocateBufferCUDA(sphx_ParamCarrier* params);
void IndexMPMSortCUDA(void);
void MpmParticleToGrid_APIC_CUDA(void);
void MpmParticleStress_CUDA(void);
void MpmNodeUpdate_CUDA(void);
void MpmParticleUpdate_APIC_CUDA(void);
void MpmParticleToGrid_CUDA(void);
void MpmParticleUpdate_CUDA(void);
void MpmGetMomentumCUDA(void);                        /* host.cu:257 */
void MpmColorTestCUDA(void);                          /* host.cu:252 */
/* host.cu:442: every kernel of it is commented out in the reference; it
 * only synchronises. */
void ComputeMpmForce(void);

/* SetupMPMGrid (fluid_system.cpp:1004-1025): node grid over the soft
 * bounds +- 10 with cell 1.5 * pSpacing. */
void sphx_setup_mpm_grid(sphx_ParamCarrier* p, float pSpacing);
/* SetDruckerPragerParam (fluid_system.cpp:882-892): a_phi, a_psi (degrees)
 * and k_c (cohesion) to the Drucker-Prager constants, in place. */
void sphx_set_drucker_prager(sphx_ParamCarrier* p);

/* ------------------------------------------------------------------------ */
/* Batched ensemble.  No counterpart in the reference (single scene).        */
/* Scenes sit in disjoint boxes of one grid and share every launch; the     */
/* scene index is bits 8-23 of the displayPack type word (bits 0-7 keep the */
/* reference's type).                                                        */
/* ------------------------------------------------------------------------ */

/* Per-scene convergence of the pressure solve for nscenes (<= 256) scenes
 * (0: one scene); the solve ends when every scene has converged, and a
 * converged scene's pressures stay fixed meanwhile. */
int sphx_ensemble_setup(sphx_context* ctx, int nscenes);
/* Iterations and final error of each scene's last solve. */
int sphx_scene_stats(sphx_context* ctx, int* iters, double* err);

/* The two-phase step (sphx_mf_step) as HIP graphs, per buffer parity:
 * graph A = sort .. advection prediction, the learned start and `budget`
 * pressure iterations (early exit on the device); graph B = integration.
 * Between them the host reads the done flag and finishes an open solve
 * eagerly (and falls back from a learned start that missed the tolerance).
 * Build after one eager sphx_mf_step; rebuild after uploads or parameter
 * changes.  Recording and the online refresh of the learned start are not
 * part of the graphs.  With the persistent solve on, graph A ends before
 * the solve, the step launches the persistent solve (and the learned
 * start's device-side fallback) between the graphs, and a step called with
 * iters = err = NULL reads nothing back: no host round trip. */
int sphx_graph_build(sphx_context* ctx, int budget);
int sphx_graph_step(sphx_context* ctx, int* iters, double* err);
int sphx_graph_release(sphx_context* ctx);
/* Persistent solve (no counterpart in the reference, whose PressureSolve
 * makes host round trips every iteration, fluid_system_host.cu:525-552):
 * the pressure solve of every step as one cooperative launch of
 * co-resident workgroups, each iteration's DP pass, Jacobi pass and FP64
 * convergence test separated by grid barriers; with a learned start, the
 * fallback to the plain start runs on the device too.  Same arithmetic
 * and results as the per-iteration launches.  Releases captured graphs. */
int sphx_persistent_solve(sphx_context* ctx, int on);
/* {iterations, final error} of the last min(solves, max_solves) persistent
 * solves, oldest first (a device ring of 4096, restarted when the
 * persistent solve is switched on); *count = how many were written. */
int sphx_solve_history(sphx_context* ctx, int* iters, double* err, int max_solves, int* count);

/* Project autotuner.  Workgroup size per neighbour-loop kernel class (timer
 * names iisph_density, iisph_dii, iisph_aii, iisph_dp, iisph_iter,
 * mf_drift, mf_alpha, mf_force, mlp_features, mlp_infer): 64, 128, 192 or
 * 256.  (The list build stages 256-particle tiles in LDS, and the a_ii pass
 * that evaluates the learned start shares one LDS copy of the weights
 * among four waves: both fixed at 256.)  sphx_autotune times each class at 64, 128 and 256 over
 * `steps` real two-phase steps per size and keeps the fastest; step graphs
 * are recaptured.  Sizes never change a particle's arithmetic. */
int sphx_tune_set(sphx_context* ctx, const char* kernel, int block);
int sphx_tune_get(sphx_context* ctx, const char* kernel, int* block);
int sphx_autotune(sphx_context* ctx, int steps);
/* The autotuner's cache: a JSON file of entries {"arch", "bucket", "kernel",
 * "block"}, keyed by (GPU architecture, size bucket = floor(log2 of the
 * context's particle count), kernel).  _save merges the context's current
 * sizes into the file under its own key (other entries are kept); _load
 * applies the entries for the context's architecture and bucket, writing
 * how many it applied to *applied (0 when the file is absent).  No
 * counterpart in the reference (fixed block size, fluid_system_host.cu:108). */
int sphx_tune_cache_save(sphx_context* ctx, const char* path);
int sphx_tune_cache_load(sphx_context* ctx, const char* path, int* applied);

/* ------------------------------------------------------------------------ */
/* Stokesian mobility (SPH_Solver/stokesian/Mobility.cu)                     */
/* ------------------------------------------------------------------------ */

/* Polydisperse Rotne-Prager-Yamakawa mobility applied on MFMA:
 * U = M F for F, U [3n][nrhs] (row 3p + a), M_ii = I/(6 pi mu a_i),
 * M_ij = (1/(8 pi mu r)) [(1 + (a_i^2+a_j^2)/(3r^2)) I + (1 - (a_i^2+a_j^2)/r^2) rr]
 * (non-overlapping spheres); its equal-radius limit times 6 pi mu a is
 * PairwiseMobMatrix's A1N (:37-82).  No counterpart in the reference for
 * several right-hand sides or unequal radii. */
typedef struct sphx_rpy sphx_rpy;
int sphx_rpy_create(sphx_rpy** out, int n, int nrhs);   /* nrhs 16, 32, 48, 64 */
int sphx_rpy_destroy(sphx_rpy* h);
int sphx_rpy_set_spheres(sphx_rpy* h, const float* pos, const float* radius, float mu);
/* Reduction segments per 64-sphere row tile (autotuned; default 16). */
int sphx_rpy_set_split(sphx_rpy* h, int split);
/* mode 1: bf16; 3: bf16x3 (hi*hi + hi*lo + lo*hi); 6: bf16x3 plus the
 * second-order split terms.  Host arrays, synchronous. */
int sphx_rpy_apply(sphx_rpy* h, const float* F, float* U, int mode);
/* Same on device arrays, asynchronous on the handle's stream. */
int sphx_rpy_apply_device(sphx_rpy* h, const float* dF, float* dU, int mode);
int sphx_rpy_sync(sphx_rpy* h);
/* Time each tile-kernel launch with HIP events on the handle's stream
 * (benchmarks); sphx_rpy_timing_get returns the total ms and count since
 * timing was switched on and forgets them. */
int sphx_rpy_timing(sphx_rpy* h, int on);
int sphx_rpy_timing_get(sphx_rpy* h, double* ms, int* count);
/* The reference walkthrough step on device buffers (equal spheres a = 1,
 * units 1/(6 pi mu a)): near-field force added to cuForce when with_force
 * (Kern_getForce :257-300), background velocity interpolated from the MPM
 * nodes of `nodes` or zero when NULL (Kern_interpolateVelocity :322-366),
 * U = U_bg + F + sum A1N F and Omega = sum B1N F, x += U dt
 * (Kern_getmobu_walkthrough :207-255; positions double-buffered, A10). */
int sphx_rpy_walkthrough(sphx_context* nodes, const sphx_stokesianBufList* buf, int n,
                         int with_force);
/* No counterpart in the reference: sphx_rpy_walkthrough for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results; the position double buffer is a static device array of
 * SPHX_WALK_NOALLOC_MAX spheres (16 B each, 1 MiB) per GPU, held until the
 * process ends, so it calls no hipMalloc/hipFree.  n above the maximum is
 * an error.  Calls that overlap on one GPU share that array and race. */
#define SPHX_WALK_NOALLOC_MAX 65536
int sphx_rpy_walkthrough_noalloc(sphx_context* nodes, const sphx_stokesianBufList* buf, int n,
                                 int with_force);
/* MVmultiply / getMobU (:85-109): u = mat f, mat (6n x 6n) row-major,
 * device pointers. */
int sphx_mob_gemv(const float* mat, const float* f, float* u, int nsize);

/* Legacy mobility entry points (Mobility.cuh:19-29; C aliases of the C++
 * declarations, which are exported under their C++ names as well). */
void setSignature(void);
void getMobU(float* mat, float* f, float* u, int nsize);
void getMobU_walkthrough(sphx_stokesianBufList buflist, int pnum);

/* ------------------------------------------------------------------------ */
/* Learned warm start and drift closure.  No counterpart in the reference:  */
/* a 24-64-64-4 ReLU MLP (6,020 weights; layout W1[24][64] b1[64]           */
/* W2[64][64] b2[64] W3[64][4] b3[4]) evaluated per fluid particle after    */
/* the a_ii pass.  Output 0 replaces the warm start p0 = 0.5 p(t-1) of      */
/* fluid_system_kern.cu:899 by p0 = max(0, p(t-1) + p_scale*y0/10); outputs */
/* 1-3 predict the change of the phase-1 drift (multiphase.cu:131-163)      */
/* under the pressure update and are added to the next step's drift.  A     */
/* solve from the learned start that misses options.tol within max_iters is */
/* redone from the plain start (iterations of both are reported).           */
/* ------------------------------------------------------------------------ */
int sphx_mlp_nparam(void);
/* He-initialised hidden layers, zero output layer (predicts no change);
 * p_scale (Pa) and u_scale (m/s) normalise pressures and drift. */
int sphx_mlp_init(sphx_context* ctx, uint64_t seed, float p_scale, float u_scale);
int sphx_mlp_set_weights(sphx_context* ctx, const float* w, int n);
int sphx_mlp_get_weights(sphx_context* ctx, float* w, int n);
/* Flags: learned warm start, learned drift correction, sample recording
 * (drift targets are recorded only while the drift correction is off). */
int sphx_mlp_mode(sphx_context* ctx, int warm, int drift, int record);
/* Online refresh: every `every` solves (0 = off) run `iters` Adam steps of
 * `batch` samples drawn on the device from the sample store, which then
 * records every solve as a ring; lw[4] weights the outputs' errors (with
 * lw[1..3] zero the recorder skips the extra drift pass of the drift
 * targets). */
int sphx_mlp_online(sphx_context* ctx, int every, int iters, int batch, float lr,
                    const float* lw);
/* on != 0: a refresh's Adam steps run on a second stream (default priority)
 * while the next solve uses the current weights; the solve after that
 * adopts the refreshed ones.  Calls that read or set weights or samples,
 * graph builds and graph steps first wait for a running refresh.  No
 * counterpart in the reference (the learned start is new). */
int sphx_mlp_online_async(sphx_context* ctx, int on);
/* Recording keeps every `stride`-th particle slot of a solve (the first
 * slot rotating with the solve count; stride 1 = all); the online ring then
 * holds as many solves as the sample store holds full solves.  No
 * counterpart in the reference. */
int sphx_mlp_record_stride(sphx_context* ctx, int stride);
int sphx_mlp_set_drift_blend(sphx_context* ctx, float blend);
/* Sample store: one sample per particle slot and solve (masked when the
 * slot is boundary or out of the grid). */
int sphx_mlp_reserve(sphx_context* ctx, int64_t max_samples);
int sphx_mlp_clear_samples(sphx_context* ctx);
int sphx_mlp_samples(sphx_context* ctx, int64_t* total, int64_t* valid);
int sphx_mlp_download_samples(sphx_context* ctx, int64_t start, int64_t count,
                              float* X, float* T, float* M);
/* Features [num][24] of the current state (after the a_ii pass), the
 * previous pressures [num] and each particle's pressure increment of the
 * last solve [num] (feature 20); any may be NULL. */
int sphx_mlp_features(sphx_context* ctx, float* X, float* pprev, float* dp_last);
/* The feature buffer as the last pass left it: the a_ii pass fills it when
 * the network is in use (warm start or recording); with the warm start on
 * and a recording stride above 1 it fills only the rows the recorder keeps
 * (the others hold an earlier step's).  sphx_mlp_features recomputes every
 * row. */
int sphx_mlp_download_features(sphx_context* ctx, float* X, float* pprev);
/* The network with FP16 weights and FP32 accumulation: Y [n][4]. */
int sphx_mlp_eval(sphx_context* ctx, const float* X, int n, float* Y);
/* Loss (1/n_valid) sum M_s sum_o lw_o (y_so - T_so)^2 and its gradient
 * (FP32 master weights) by the training kernels; M may be NULL. */
int sphx_mlp_grad(sphx_context* ctx, const float* X, const float* T, const float* M,
                  int n, const float* lw, float* grad, double* loss);
/* Adam (beta 0.9/0.999) over the recorded valid samples; loss_out[epochs]
 * may be NULL. */
int sphx_mlp_train(sphx_context* ctx, int epochs, int batch, float lr, const float* lw,
                   uint64_t seed, double* loss_out);
int sphx_mlp_stats(sphx_context* ctx, int64_t* learned, int64_t* fallbacks,
                   int* last_fallback);
/* Guard of the learned start against a broken prediction (single scene;
 * factor 10 by default, 0 switches it off): when a learned solve's first
 * iteration leaves an error above factor * tol, one probe iteration from
 * the plain start p0 = warm_start * p(t-1) runs and the solve continues
 * from the better first iterate; the probe counts as an iteration.
 * Decided on the device (graph-safe).  A factor near 1 makes the choice
 * greedy per step, which prefers the plain start's slowly growing
 * compression and loses the learned start's pressure over later steps. */
int sphx_mlp_guard(sphx_context* ctx, float factor);
/* Probes run and probes that kept the plain start, since sphx_mlp_init. */
int sphx_mlp_guard_stats(sphx_context* ctx, int64_t* probes, int64_t* plain_picks);

/* ------------------------------------------------------------------------ */
/* Kernel timing (HIP events on the context stream).                         */
/* ------------------------------------------------------------------------ */
int sphx_timing_enable(sphx_context* ctx, int on);
/* Time only the named kernel (null or "": every kernel). */
int sphx_timing_only(sphx_context* ctx, const char* kernel);
int sphx_timing_reset(sphx_context* ctx);
/* Accumulated device time (ms) and launch count of the named kernel. */
int sphx_timing_get(sphx_context* ctx, const char* kernel, double* total_ms,
                    int* launches);

/* ------------------------------------------------------------------------ */
/* Synthetic scenes (host only; no GPU needed).                              */
/* ------------------------------------------------------------------------ */

/* FluidSystem::SetupGrid (fluid_system.cpp:119-145): from volmin, volmax,
 * cellsize and simscale fill gridmin/max/size/res/total, gridIdfac and
 * neighborid with the reference's float/double expression types. */
void sphx_setup_grid(sphx_ParamCarrier* p);
/* Kernel constants of fluid_system.cpp:308-313 from smoothradius
 * (cellsize = 2h, kpoly6, kspiky, kspikydiff, klaplacian, kspline). */
void sphx_setup_kernels(sphx_ParamCarrier* p);

/* ------------------------------------------------------------------------ */
/* State I/O in the reference's formats (host only).                        */
/* ------------------------------------------------------------------------ */
/* The values FluidSystem::ParseXML reads (fluid_system.cpp:897-977) that
 * are not ParamCarrier fields. */
typedef struct sphx_SceneXml {
  int maxPointNum;          /* Num */
  float scaleP;             /* ScaleP */
  int fluidnum;             /* FluidCount */
  sphx_float3 fvmin[2];     /* VolMin0, VolMin1 */
  sphx_float3 fvmax[2];     /* VolMax0, VolMax1 */
  float massratio[3];       /* MassRatio "a,b,c" */
  float densratio[3];       /* DensityRatio */
  float viscratio[3];       /* ViscRatio */
} sphx_SceneXml;
/* scene.xml text (Fluid, MultiScene id=caseid, BoundInfo): fills the
 * ParamCarrier fields ParseXML sets, with the Drucker-Prager constants
 * converted from degrees (SetDruckerPragerParam, fluid_system.cpp:882-892);
 * absent values read as 0 (catXMLHelper.cpp:17-25); scene may be NULL.
 * Returns -1 (message in sphx_io_last_error) when Fluid, MultiScene or the
 * case is missing. */
int sphx_scene_xml_parse(const char* text, int caseid, sphx_ParamCarrier* params,
                         sphx_SceneXml* scene);
int sphx_scene_xml_load(const char* path, int caseid, sphx_ParamCarrier* params,
                        sphx_SceneXml* scene);
/* Text frame of FluidSystem::outputFile (FluidIOUtility.cpp:17-38): the
 * count, then "x y z mass restdens type" per particle (%f, %d).  Reading
 * fills those fields of the first min(count, cap) particles; *n = count. */
int sphx_frame_write(const char* path, const sphx_displayPack* disp,
                     const sphx_calculationPack* calc, int n);
int sphx_frame_read(const char* path, sphx_displayPack* disp, sphx_calculationPack* calc,
                    int cap, int* n);
/* Checkpoint: the ParamCarrier and the displayPack / calculationPack arrays
 * in the reference's layout (binary, with their sizes checked on reading).
 * read copies the particles when cap >= *n (cap = 0: size query). */
int sphx_checkpoint_write(const char* path, const sphx_ParamCarrier* params,
                          const sphx_displayPack* disp, const sphx_calculationPack* calc,
                          int n);
int sphx_checkpoint_read(const char* path, sphx_ParamCarrier* params, sphx_displayPack* disp,
                         sphx_calculationPack* calc, int cap, int* n);
/* A context's parameters and particles (slot order) to a checkpoint, and
 * back into a context created for at least that many particles. */
int sphx_checkpoint_save(sphx_context* ctx, const char* path);
int sphx_checkpoint_restore(sphx_context* ctx, const char* path);
/* Boundary particles from a shape file (LoadBoundary, AddShape.cpp:265-359):
 * "WALL min max" and "OPEN_BOX min max thickness" commands, each wall a
 * lattice of spacing (bmass / bRestdensity)^(1/3) / simscale as
 * SetupAddWall fills it.  Particles go to slots start .. cap-1 (either
 * array may be NULL to count); *added = the number the file describes. */
int sphx_boundary_parse(const char* text, const sphx_ParamCarrier* params,
                        sphx_displayPack* disp, sphx_calculationPack* calc, int start,
                        int cap, int* added);
int sphx_boundary_load(const char* path, const sphx_ParamCarrier* params,
                       sphx_displayPack* disp, sphx_calculationPack* calc, int start,
                       int cap, int* added);
const char* sphx_io_last_error(void);

/* Counts for a scene so callers can size arrays; returns total particles. */
typedef struct sphx_SceneInfo {
  int n_fluid;
  int n_boundary;
  int n_total;
  int n_phase;
  float spacing;
} sphx_SceneInfo;

/* Single-phase dam break: nx*ny*nz fluid lattice of spacing s in the corner
 * of a box [0,Lx]x[0,Ly]x[0,Lz] (metres) with a boundary-particle box of
 * `layers` layers; jitter +-jitter from seed.  kernel_family as in
 * sphx_SolverOptions.wcsph_kernel.  Fills params (grid, constants) and the
 * particle arrays (may be NULL to query sizes). */
int sphx_scene_dambreak(int nx, int ny, int nz, float s, float Lx, float Ly,
                        float Lz, float h, int layers, float jitter,
                        uint64_t seed, int kernel_family,
                        sphx_ParamCarrier* params, sphx_displayPack* disp,
                        sphx_calculationPack* calc, sphx_SceneInfo* info);

/* Two-phase bubble tank (water 1000 / air 1.2): n_fluid particles on a
 * lattice of spacing s = (fill*Lx*Ly*Lz/n_fluid)^(1/3) filling a Lx*Ly*Lz tank
 * from the floor (y-major order); n_bubbles seeded spheres, radii uniform in
 * [rmin, rmax], mark air.  Every fluid particle carries the solver mass
 * 1000*s^3 and rest density 1000; `alpha` receives [n_total][2] volume
 * fractions (water, air).  Boundary: `layers` wall layers, open top. */
int sphx_scene_bubbles(int n_fluid, float Lx, float Ly, float Lz, float fill, float h,
                       int n_bubbles, float rmin, float rmax, int layers, uint64_t seed,
                       sphx_ParamCarrier* params, sphx_displayPack* disp,
                       sphx_calculationPack* calc, float* alpha, sphx_SceneInfo* info);

/* Granular block for the MPM transfer: n particles uniform in [lo,hi]^3
 * (metres) over a res^3 node grid of spacing dx from the origin; mass
 * rho (hi-lo)^3 / n; G, K from E and nu (fluid_system.cpp:290-291);
 * Drucker-Prager friction angle phi_deg, no cohesion or dilation;
 * deformation gradients I + N(0, f_noise^2); soft walls 3 nodes inside the
 * grid; SPH sort grid of cells 4 dx.  Returns n (arrays may be NULL). */
int sphx_scene_sand(int n, float lo, float hi, float dx, int res, float E, float nu,
                    float phi_deg, float rho, float f_noise, uint64_t seed,
                    sphx_ParamCarrier* params, sphx_displayPack* disp,
                    sphx_calculationPack* calc, sphx_SceneInfo* info);

/* Batched ensemble: nscenes two-phase tanks of n_fluid[s] particles at the
 * spacing of the 2^20-particle tank, bubbles seeded per scene, in
 * cell-aligned boxes with 2-cell gaps of one grid; the scene index goes into
 * bits 8-23 of the type word; scene_first [nscenes + 1].  Returns the
 * particle count (arrays may be NULL). */
int sphx_scene_ensemble(int nscenes, const int* n_fluid, float fill, float h,
                        float bubbles_per_million, float rmin, float rmax, int layers,
                        uint64_t seed, sphx_ParamCarrier* params, sphx_displayPack* disp,
                        sphx_calculationPack* calc, float* alpha, int* scene_first,
                        sphx_SceneInfo* info);

#ifdef __cplusplus
}
#endif

#endif /* SPHX_H */
