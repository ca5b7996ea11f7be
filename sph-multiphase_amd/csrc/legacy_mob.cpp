// The mobility entry points under the C++ names the reference declares
// (SPH_Solver/stokesian/Mobility.cuh:19-29), forwarding to the C aliases.
// Kept apart from sphx.h, whose C declarations share these signatures.
struct stokesianBufList {   // Mobility.cuh:6-16, layout of sphx_stokesianBufList
  void* cuUnew;
  void* dispBuffer;
  void* cuOmega;
  void* cuForce;
  float dt;
  float simscale;
};

extern "C" {
struct sphx_stokesianBufList;
void sphx_legacy_setsignature(void);
void sphx_legacy_getmobu(float* mat, float* f, float* u, int nsize);
void sphx_legacy_walkthrough(const void* buflist, int pnum);
}

void setSignature() { sphx_legacy_setsignature(); }
void getMobU(float* mat, float* f, float* u, int nsize) { sphx_legacy_getmobu(mat, f, u, nsize); }
void getMobU_walkthrough(stokesianBufList buflist, int pnum) {
  sphx_legacy_walkthrough(&buflist, pnum);
}
