// Legacy drop-in entry points (fluid_system_host.cuh:44-107 and the C++
// linkage CarryParam of fluid_system_kern.cuh:71).  One process-wide default
// context, a device synchronisation after every call, errors printed and
// latched as the reference's cudaGetLastError() checks do
// (fluid_system_host.cu:154-157 and alike).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "../../include/sphx.h"

namespace {

sphx_context* g_ctx = nullptr;
sphx_ParamCarrier g_params;
bool g_have_params = false;

void report(const char* where, int rc) {
  if (rc != 0) std::fprintf(stderr, "SPHX ERROR: %s: %s\n", where, sphx_last_error());
}

void run(const char* where, int (*fn)(sphx_context*)) {
  if (!g_ctx) {
    std::fprintf(stderr, "SPHX ERROR: %s: FluidSetupCUDA has not been called\n", where);
    return;
  }
  int rc = fn(g_ctx);
  if (rc == 0) rc = sphx_sync(g_ctx);
  report(where, rc);
}

}  // namespace

extern "C" {

sphx_context* sphx_legacy_context(void) { return g_ctx; }

void cudaInit(int, char**) {
  int n = sphx_device_count();
  hipDeviceProp_t p;
  if (n > 0 && hipGetDeviceProperties(&p, 0) == hipSuccess) {
    std::printf("-- HIP --\n");
    std::printf("Name:       %s\n", p.name);
    std::printf("Arch:       %s\n", p.gcnArchName);
    std::printf("Global Mem: %zu\n", (size_t)p.totalGlobalMem);
    std::printf("Shared/Blk: %zu\n", (size_t)p.sharedMemPerBlock);
    std::printf("Warp Size:  %d\n", p.warpSize);
    std::printf("Thrds/Blk:  %d\n", p.maxThreadsPerBlock);
    std::printf("CUs:        %d\n", p.multiProcessorCount);
  } else {
    std::fprintf(stderr, "SPHX ERROR: cudaInit: no HIP device\n");
  }
}

void cudaExit(int, char**) {}

void FluidClearCUDA(void) {
  if (g_ctx) sphx_destroy(g_ctx);
  g_ctx = nullptr;
}

void FluidSetupCUDA(sphx_ParamCarrier* params) {
  if (!params) return;
  FluidClearCUDA();
  g_params = *params;
  g_have_params = true;
  int rc = sphx_create(&g_ctx, params);
  report("FluidSetupCUDA", rc);
}

void CarryParam(sphx_ParamCarrier* hostCarrier) {
  if (!hostCarrier) return;
  g_params = *hostCarrier;
  g_have_params = true;
  if (g_ctx) report("CarryParam", sphx_set_params(g_ctx, hostCarrier));
}

void computeNumBlocks(int numPnts, int maxThreads, int* numBlocks, int* numThreads) {
  int t = numPnts < maxThreads ? numPnts : maxThreads;
  int b = (t == 0) ? 1 : (numPnts % t != 0 ? numPnts / t + 1 : numPnts / t);
  if (t == 0) t = 1;
  if (numBlocks) *numBlocks = b;
  if (numThreads) *numThreads = t;
}

void GetParticleIndexCUDA(void) { run("GetParticleIndexCUDA", sphx_grid_hash); }
void GetGridListCUDA(void) { run("GetGridListCUDA", sphx_grid_sort); }
void RearrageDataCUDA(void) { run("RearrageDataCUDA", sphx_grid_reorder); }
void ComputeBoundaryDensity(void) { run("ComputeBoundaryDensity", sphx_boundary_volume); }
void MfComputePressureCUDA(void) { run("MfComputePressureCUDA", sphx_wcsph_density_pressure); }
void MfComputeForceCUDA(void) { run("MfComputeForceCUDA", sphx_wcsph_force); }
void MfAdvanceCUDA(void) { run("MfAdvanceCUDA", sphx_wcsph_advance); }
void ComputeSolidTensorCUDA(void) { run("ComputeSolidTensorCUDA", sphx_solid_tensor); }
void ComputeSolidForceCUDA(void) { run("ComputeSolidForceCUDA", sphx_solid_force); }
// ComputeSolidTensor_X / ComputeSolidForce_X are empty kernels in the
// reference (sph_solid.cu:650-661): nothing to do
void ComputeSolidTensorX_CUDA(void) {}
void ComputeSolidForceX_CUDA(void) {}
void InitializeSolid_CUDA(void) { run("InitializeSolid_CUDA", sphx_grid_build); }
void ComputeDensityIISPH_CUDA(void) { run("ComputeDensityIISPH_CUDA", sphx_iisph_density); }
void PredictAdvection(void) { run("PredictAdvection", sphx_iisph_predict_advection); }

void PressureSolve(void) {
  if (!g_ctx) {
    std::fprintf(stderr, "SPHX ERROR: PressureSolve: FluidSetupCUDA has not been called\n");
    return;
  }
  int iters = 0;
  double err = 0.0;
  int rc = sphx_iisph_pressure_solve(g_ctx, &iters, &err);
  report("PressureSolve", rc);
  if (rc == 0 && std::getenv("SPHX_VERBOSE"))
    std::printf("pressure solve: %d iterations, mean |rho-rho0|/rho0 %g\n", iters, err);
}

void Integration(void) { run("Integration", sphx_iisph_integrate); }
void MfComputeDriftVelCUDA(void) { run("MfComputeDriftVelCUDA", sphx_mf_drift_velocity); }
void MfComputeAlphaAdvanceCUDA(void) { run("MfComputeAlphaAdvanceCUDA", sphx_mf_alpha_advance); }
void MfComputeCorrectionCUDA(void) { run("MfComputeCorrectionCUDA", sphx_mf_correction); }
void ComputeForceCUDA_ProjectU(void) { run("ComputeForceCUDA_ProjectU", sphx_mf_force); }

// MPM (fluid_system_host.cu:228-296)
void MpmAllocateBufferCUDA(sphx_ParamCarrier* params) {
  if (!g_ctx) {
    std::fprintf(stderr, "SPHX ERROR: MpmAllocateBufferCUDA: FluidSetupCUDA has not been called\n");
    return;
  }
  if (!params) {
    std::fprintf(stderr, "SPHX ERROR: MpmAllocateBufferCUDA: null ParamCarrier\n");
    return;
  }
  int rc = sphx_set_params(g_ctx, params);
  if (rc == 0) {
    g_params = *params;
    rc = sphx_mpm_setup(g_ctx, 0, 0);
  }
  report("MpmAllocateBufferCUDA", rc);
}
// initMpm (MpmSolver.cu:39-73) stores node positions and SPH cells; node
// positions are computed where they are used here, so this only checks
// that the node grid exists.
void IndexMPMSortCUDA(void) {
  run("IndexMPMSortCUDA", [](sphx_context* c) { int n; return sphx_mpm_active_blocks(c, &n); });
}
void MpmParticleToGrid_APIC_CUDA(void) {
  run("MpmParticleToGrid_APIC_CUDA", [](sphx_context* c) { return sphx_mpm_p2g(c, 1); });
}
void MpmParticleToGrid_CUDA(void) {
  run("MpmParticleToGrid_CUDA", [](sphx_context* c) { return sphx_mpm_p2g(c, 0); });
}
void MpmParticleStress_CUDA(void) { run("MpmParticleStress_CUDA", sphx_mpm_stress); }
void MpmNodeUpdate_CUDA(void) { run("MpmNodeUpdate_CUDA", sphx_mpm_node_update); }
void MpmParticleUpdate_APIC_CUDA(void) {
  run("MpmParticleUpdate_APIC_CUDA", [](sphx_context* c) { return sphx_mpm_g2p(c, 1); });
}
void MpmParticleUpdate_CUDA(void) {
  run("MpmParticleUpdate_CUDA", [](sphx_context* c) { return sphx_mpm_g2p(c, 0); });
}
void MpmGetMomentumCUDA(void) { run("MpmGetMomentumCUDA", sphx_mpm_get_momentum); }
void MpmColorTestCUDA(void) { run("MpmColorTestCUDA", sphx_mpm_color_test); }
void ComputeMpmForce(void) { run("ComputeMpmForce", [](sphx_context*) { return 0; }); }

// Mobility (Mobility.cuh:19-29).  The antisymmetric tensor of
// setSignature (:20-35) is written out inside the kernels.
void setSignature(void) {}
void getMobU(float* mat, float* f, float* u, int nsize) {
  report("getMobU", sphx_mob_gemv(mat, f, u, nsize));
}
void getMobU_walkthrough(sphx_stokesianBufList buflist, int pnum) {
  report("getMobU_walkthrough", sphx_rpy_walkthrough(g_ctx, &buflist, pnum, 1));
}

void sphx_legacy_setsignature(void) { setSignature(); }
void sphx_legacy_getmobu(float* mat, float* f, float* u, int nsize) { getMobU(mat, f, u, nsize); }
void sphx_legacy_walkthrough(const void* buflist, int pnum) {
  getMobU_walkthrough(*static_cast<const sphx_stokesianBufList*>(buflist), pnum);
}

void sphx_legacy_upload(const sphx_displayPack* disp, const sphx_calculationPack* calc, int start,
                        int end) {
  if (!g_ctx) {
    std::fprintf(stderr, "SPHX ERROR: upload: FluidSetupCUDA has not been called\n");
    return;
  }
  report("upload", sphx_upload(g_ctx, disp + start, calc + start, start, end - start, 1));
}

void sphx_legacy_download(sphx_displayPack* disp, sphx_calculationPack* calc, int count) {
  if (!g_ctx) return;
  report("download", sphx_download(g_ctx, disp, calc, 0, count));
}

}  // extern "C"

// The reference declares CarryParam with C++ linkage and a reference
// argument (fluid_system_kern.cuh:71).  The mangled name depends only on the
// class name, so an incomplete declaration is enough to export the symbol
// the reference's object files expect.
struct ParamCarrier;
void CarryParam(ParamCarrier& p) { CarryParam(reinterpret_cast<sphx_ParamCarrier*>(&p)); }
