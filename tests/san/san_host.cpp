// Host-side code of the library (csrc/io.cpp, csrc/scenes.cpp) and the
// oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every parser
// and scene generator on valid inputs, on every truncation of them and on
// byte-flipped variants; file readers on truncated and garbled files.
// Built and run by tests/test_sanitizers.py (CPU only).  Exit status 0
// means no sanitizer report (the sanitizers abort on the first one).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/sphx.h"

namespace {

const char* kXml =
    "<?xml version=\"1.0\"?>\n"
    "<!-- a scene file in the reference's layout -->\n"
    "<Fluid>\n"
    "  <AccelLimit> 1500.0 </AccelLimit><VelLimit> 100 </VelLimit>\n"
    "  <Gravity> 0,-9.8,0.5 </Gravity><VolMin> -60, -20, -60 </VolMin>\n"
    "  <VolMax> 60, 200, 60 </VolMax>\n"
    "</Fluid>\n"
    "<MultiScene id=\"1\">\n"
    "  <Num> 500000 </Num><SimScale> 0.005 </SimScale><DT> 0.00005 </DT>\n"
    "  <Radius> 0.01 </Radius><SmoothRadius> 0.015 </SmoothRadius>\n"
    "  <Viscosity> 1.5 </Viscosity><Mass> 0.00010543 </Mass><RestDensity> 1000 </RestDensity>\n"
    "  <FluidCount> 3 </FluidCount><SoftBoundMin> -40, 0, -40 </SoftBoundMin>\n"
    "  <SoftBoundMax> 40, 100, 40 </SoftBoundMax><VolMin0> -20, 2, -20 </VolMin0>\n"
    "  <VolMax0> 20, 10, 20 </VolMax0><!--VolMin0> -7, 2, -7 </VolMin0-->\n"
    "  <MassRatio> 1., 2., 3. </MassRatio><DensityRatio> 1., 0.5, 0.25 </DensityRatio>\n"
    "  <ViscRatio> 1, 4, 9 </ViscRatio><SolidK> 1000000000 </SolidK>\n"
    "  <DP_Phi> 15 </DP_Phi><DP_Psi> 10 </DP_Psi><DP_cohesion> 10.0 </DP_cohesion>\n"
    "</MultiScene>\n"
    "<BoundInfo id=\"1\">\n"
    "  <Viscosity> 100.0 </Viscosity><RestDensity> 600.0 </RestDensity>\n"
    "  <Mass> 0.00010543 </Mass>\n"
    "</BoundInfo>\n";

const char* kShape =
    "WALL -1,0,-1 1,0.05,1\n"
    "OPEN_BOX -0.5,0,-0.5 0.5,0.4,0.5 0.03\n";

int g_checks = 0;

void check(bool ok, const char* what) {
  ++g_checks;
  if (!ok) {
    std::fprintf(stderr, "san_host: check failed: %s\n", what);
    std::exit(2);
  }
}

// Every prefix of `text`, then byte flips: the parser must return, with or
// without an error, and never read or write out of bounds.
template <class F>
void mutate(const std::string& text, F&& parse, int flips, unsigned seed) {
  for (size_t n = 0; n <= text.size(); ++n) parse(text.substr(0, n));
  std::mt19937 rng(seed);
  const char junk[] = "<>/=\",.-0123456789eE \n\tWALLOPEN_BOX";
  for (int t = 0; t < flips; ++t) {
    std::string s = text;
    const int k = 1 + (int)(rng() % 4);
    for (int q = 0; q < k; ++q) {
      size_t at = rng() % s.size();
      s[at] = (rng() % 3 == 0) ? (char)(rng() % 256) : junk[rng() % (sizeof(junk) - 1)];
    }
    parse(s);
  }
}

void test_xml() {
  sphx_ParamCarrier p;
  sphx_SceneXml sc;
  std::memset(&p, 0, sizeof(p));
  check(sphx_scene_xml_parse(kXml, 1, &p, &sc) == 0, "scene.xml parses");
  check(std::fabs(p.smoothradius - 0.015f) < 1e-7f && sc.fluidnum == 3, "scene.xml values");
  check(sphx_scene_xml_parse(kXml, 7, &p, nullptr) != 0, "missing case is an error");
  mutate(kXml, [](const std::string& s) {
    sphx_ParamCarrier q;
    sphx_SceneXml x;
    (void)sphx_scene_xml_parse(s.c_str(), 1, &q, &x);
  }, 4000, 1);
}

void test_boundary() {
  sphx_ParamCarrier p;
  std::memset(&p, 0, sizeof(p));
  p.bmass = 0.001f;
  p.bRestdensity = 1000.f;
  p.simscale = 1.f;
  int added = 0;
  check(sphx_boundary_parse(kShape, &p, nullptr, nullptr, 0, 0, &added) == 0 && added > 0,
        "shape file counts");
  std::vector<sphx_displayPack> d((size_t)added);
  std::vector<sphx_calculationPack> c((size_t)added);
  int again = 0;
  check(sphx_boundary_parse(kShape, &p, d.data(), c.data(), 0, added, &again) == 0 &&
            again == added, "shape file fills");
  // half the capacity: the rest is counted, nothing written past cap
  std::vector<sphx_displayPack> h((size_t)added / 2);
  check(sphx_boundary_parse(kShape, &p, h.data(), nullptr, 0, added / 2, &again) == 0,
        "shape file with a short buffer");
  mutate(kShape, [&](const std::string& s) {
    int a = 0;
    std::vector<sphx_displayPack> dd(64);
    std::vector<sphx_calculationPack> cc(64);
    (void)sphx_boundary_parse(s.c_str(), &p, dd.data(), cc.data(), 0, 64, &a);
  }, 3000, 2);
  // spacing that a float cannot step (must be refused, not looped on)
  sphx_ParamCarrier bad = p;
  bad.bmass = 0.f;
  check(sphx_boundary_parse(kShape, &bad, nullptr, nullptr, 0, 0, &added) != 0,
        "zero spacing is an error");
}

template <class T>
void write_bytes(const char* path, const std::vector<T>& v, size_t bytes) {
  FILE* f = std::fopen(path, "wb");
  check(f != nullptr, "temp file");
  std::fwrite(v.data(), 1, bytes, f);
  std::fclose(f);
}

std::vector<char> read_bytes(const char* path) {
  std::vector<char> b;
  FILE* f = std::fopen(path, "rb");
  if (!f) return b;
  char buf[4096];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + k);
  std::fclose(f);
  return b;
}

void test_files(const char* dir) {
  sphx_ParamCarrier p;
  sphx_SceneInfo info;
  int n = sphx_scene_dambreak(4, 4, 4, 0.01f, 0.08f, 0.06f, 0.06f, 0.01f, 1, 1e-4f, 3, 1, &p,
                              nullptr, nullptr, &info);
  check(n > 0, "dam break counts");
  std::vector<sphx_displayPack> d((size_t)n);
  std::vector<sphx_calculationPack> c((size_t)n);
  check(sphx_scene_dambreak(4, 4, 4, 0.01f, 0.08f, 0.06f, 0.06f, 0.01f, 1, 1e-4f, 3, 1, &p,
                            d.data(), c.data(), &info) == n, "dam break fills");
  const std::string frame = std::string(dir) + "/frame.txt";
  const std::string ckp = std::string(dir) + "/state.ckp";
  const std::string cut = std::string(dir) + "/cut.bin";
  check(sphx_frame_write(frame.c_str(), d.data(), c.data(), n) == 0, "frame write");
  check(sphx_checkpoint_write(ckp.c_str(), &p, d.data(), c.data(), n) == 0, "checkpoint write");
  std::vector<sphx_displayPack> d2((size_t)n);
  std::vector<sphx_calculationPack> c2((size_t)n);
  int m = 0;
  check(sphx_frame_read(frame.c_str(), d2.data(), c2.data(), n, &m) == 0 && m == n, "frame read");
  sphx_ParamCarrier q;
  check(sphx_checkpoint_read(ckp.c_str(), &q, d2.data(), c2.data(), n, &m) == 0 && m == n,
        "checkpoint read");
  check(std::memcmp(d.data(), d2.data(), sizeof(sphx_displayPack) * n) == 0, "round trip");
  // every truncation and some garbling of both files; readers get a small
  // buffer, so a count they believe must not make them write past it
  std::mt19937 rng(4);
  for (const std::string* src : {&frame, &ckp}) {
    std::vector<char> b = read_bytes(src->c_str());
    check(!b.empty(), "file bytes");
    for (size_t len = 0; len <= b.size(); len += (len < 256 ? 1 : 97)) {
      write_bytes(cut.c_str(), b, len);
      std::vector<sphx_displayPack> dd(8);
      std::vector<sphx_calculationPack> cc(8);
      int k = 0;
      (void)sphx_frame_read(cut.c_str(), dd.data(), cc.data(), 8, &k);
      (void)sphx_checkpoint_read(cut.c_str(), &q, dd.data(), cc.data(), 8, &k);
    }
    for (int t = 0; t < 300; ++t) {
      std::vector<char> g = b;
      for (int r = 0; r < 3; ++r) g[rng() % g.size()] = (char)(rng() % 256);
      write_bytes(cut.c_str(), g, g.size());
      std::vector<sphx_displayPack> dd(8);
      std::vector<sphx_calculationPack> cc(8);
      int k = 0;
      (void)sphx_frame_read(cut.c_str(), dd.data(), cc.data(), 8, &k);
      (void)sphx_checkpoint_read(cut.c_str(), &q, dd.data(), cc.data(), 8, &k);
    }
  }
}

void test_scenes() {
  sphx_ParamCarrier p;
  sphx_SceneInfo info;
  // bubbles: size query, then fill
  int n = sphx_scene_bubbles(3000, 0.15f, 0.15f, 0.15f, 0.8f, 0.006f, 6, 0.01f, 0.03f, 2, 5, &p,
                             nullptr, nullptr, nullptr, &info);
  check(n > 0, "bubbles counts");
  std::vector<sphx_displayPack> d((size_t)n);
  std::vector<sphx_calculationPack> c((size_t)n);
  std::vector<float> a((size_t)n * 2);
  check(sphx_scene_bubbles(3000, 0.15f, 0.15f, 0.15f, 0.8f, 0.006f, 6, 0.01f, 0.03f, 2, 5, &p,
                           d.data(), c.data(), a.data(), &info) == n, "bubbles fills");
  n = sphx_scene_sand(500, 0.3f, 0.6f, 1.f / 32, 32, 3.5e5f, 0.3f, 30.f, 1600.f, 1e-3f, 6, &p,
                      nullptr, nullptr, &info);
  check(n == 500, "sand counts");
  d.resize((size_t)n);
  c.resize((size_t)n);
  check(sphx_scene_sand(500, 0.3f, 0.6f, 1.f / 32, 32, 3.5e5f, 0.3f, 30.f, 1600.f, 1e-3f, 6, &p,
                        d.data(), c.data(), &info) == n, "sand fills");
  const int sizes[3] = {2000, 3000, 2500};
  n = sphx_scene_ensemble(3, sizes, 0.8f, 0.01f, 512.f, 0.01f, 0.04f, 2, 7, &p, nullptr, nullptr,
                          nullptr, nullptr, &info);
  check(n > 0, "ensemble counts");
  d.resize((size_t)n);
  c.resize((size_t)n);
  a.resize((size_t)n * 2);
  std::vector<int> first(4);
  check(sphx_scene_ensemble(3, sizes, 0.8f, 0.01f, 512.f, 0.01f, 0.04f, 2, 7, &p, d.data(),
                            c.data(), a.data(), first.data(), &info) == n, "ensemble fills");
  // degenerate arguments are refused or give empty scenes, never UB
  (void)sphx_scene_dambreak(0, 3, 3, 0.01f, 0.1f, 0.1f, 0.1f, 0.01f, 1, 0.f, 1, 0, &p, nullptr,
                            nullptr, &info);
  (void)sphx_scene_bubbles(-5, 0.1f, 0.1f, 0.1f, 0.8f, 0.01f, 1, 0.01f, 0.02f, 1, 1, &p, nullptr,
                           nullptr, nullptr, &info);
  (void)sphx_scene_sand(10, 0.6f, 0.3f, 1.f / 32, 32, 3.5e5f, 0.3f, 30.f, 1600.f, 0.f, 1, &p,
                        nullptr, nullptr, &info);
  // grid and constants helpers
  sphx_ParamCarrier g;
  std::memset(&g, 0, sizeof(g));
  g.volmin = {-0.1f, 0.f, -0.1f};
  g.volmax = {0.1f, 0.3f, 0.1f};
  g.smoothradius = 0.01f;
  g.simscale = 1.f;
  sphx_setup_kernels(&g);
  sphx_setup_grid(&g);
  check(g.gridtotal == g.gridres.x * g.gridres.y * g.gridres.z && g.gridtotal > 0, "grid");
  g.softminx = {-20.f, 0.f, -20.f};
  g.softmaxx = {20.f, 30.f, 20.f};
  sphx_setup_mpm_grid(&g, 0.5f);
  check(g.mpmRes.x == 82 && g.mpmNodeNum == g.mpmRes.x * g.mpmRes.y * g.mpmRes.z, "mpm grid");
  sphx_setup_mpm_grid(&g, 0.005f);          // 2695^3 nodes: too many for an int count
  check(g.mpmNodeNum == -1, "mpm grid node count overflow is flagged");
  sphx_setup_mpm_grid(&g, 0.f);             // zero spacing: infinite counts
  check(g.mpmRes.x == 0 && g.mpmNodeNum == 0, "zero spacing gives an empty grid");
  sphx_set_drucker_prager(&g);
}

}  // namespace

int main(int argc, char** argv) {
  const char* dir = argc > 1 ? argv[1] : ".";
  test_xml();
  test_boundary();
  test_files(dir);
  test_scenes();
  std::printf("san_host: %d checks passed\n", g_checks);
  return 0;
}
