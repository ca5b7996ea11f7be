// The two-phase step as HIP graphs.
//
// A step is captured once per particle-buffer parity as two graphs:
//   A: sort, density, drift-flux stages, advection prediction, the learned
//      start (when enabled) and a fixed budget of pressure iterations;
//   B: the per-particle pressure increment of the learned start and the
//      integration.
// Between them the host reads the solver's done flag (8 bytes): a solve
// still open after the budget continues eagerly, and a learned start that
// missed the tolerance falls back to the plain start, so the plain solver
// keeps judging every step.  With the persistent solve on, that host check
// is replaced by a cooperative launch that continues a solve still open
// after the budget on the device (and returns at once when none is), so a
// step needs no host read-back.  Iterations past convergence exit on the
// device flag.  Every launch covers all scenes of a batched ensemble at once.
// Nothing in the captured work allocates, synchronises or clears memory
// with a memset node (the solver state is reset by a kernel).
#include "common.hpp"

#include <cstring>
#include <fstream>
#include <sstream>

namespace sphx {

int enqueue_solve_fixed(sphx_context* c, int n);
int iisph_density_dii(sphx_context* c);
int mf_correction_force(sphx_context* c);
int continue_solve(sphx_context* c, int* iters, double* err);
int enqueue_solve_coop(sphx_context* c, bool reset);
int read_solve(sphx_context* c, int* iters, double* err);
int solve_log_setup(sphx_context* c);
int solve_log_read(sphx_context* c, int* iters, double* err, int max, int* count);
int mlp_begin(sphx_context* c);
int mlp_fallback(sphx_context* c, int* iters, double* err);
int mlp_after(sphx_context* c, int record_ok);
int mlp_async_settle(sphx_context* c);

static void graph_release(sphx_context* c) {
  for (int p = 0; p < 2; ++p) {
    if (c->graph_a[p]) (void)hipGraphExecDestroy(c->graph_a[p]);
    if (c->graph_b[p]) (void)hipGraphExecDestroy(c->graph_b[p]);
    c->graph_a[p] = c->graph_b[p] = nullptr;
  }
  c->graph_num = -1;
}

// Host bookkeeping of graph A's stages (what the eager calls leave behind).
static void after_graph_a(sphx_context* c, int p) {
  c->cur = 1 - p;
  c->sorted_valid = true;
  c->nbr_valid = true;
  c->gh_valid = true;
  c->mf_pending = true;
}

static int stage_a(sphx_context* c, int budget) {
  SPHX_TRY(sphx_grid_build(c));
  SPHX_TRY(iisph_density_dii(c));
  SPHX_TRY(sphx_mf_drift_velocity(c));
  SPHX_TRY(sphx_mf_alpha_advance(c));
  SPHX_TRY(mf_correction_force(c));
  SPHX_TRY(sphx_iisph_predict_aii(c));
  SPHX_TRY(mlp_begin(c));
  return enqueue_solve_fixed(c, budget);
}

static int stage_b(sphx_context* c) {
  SPHX_TRY(mlp_after(c, 0));
  return sphx_iisph_integrate(c);
}

template <class F>
static int capture(sphx_context* c, F&& body, hipGraphExec_t* out) {
  SPHX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  int rc = body();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc != 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail("hipStreamEndCapture", e);
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail("hipGraphInstantiate", e);
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

// Capture the two-phase step with `budget` pressure iterations in graph A.
// The context must have run at least one eager sphx_mf_step (so every
// buffer exists); recapture after an upload or a parameter change.
int sphx_graph_build(sphx_context* c, int budget) {
  SPHX_CHECK(c, "sphx_graph_build", "null context");
  SPHX_CHECK(c->nphase > 0, "sphx_graph_build", "multiphase not set up");
  SPHX_CHECK(budget >= 1, "sphx_graph_build", "budget must be >= 1");
  SPHX_CHECK(c->mf_pending && c->nbr_cap > 0 && c->red.p, "sphx_graph_build",
             "run one sphx_mf_step before capturing");
  SPHX_TRY(ensure_device(c));
  SPHX_TRY(mlp_async_settle(c));   // no cross-stream wait inside the capture
  SPHX_HIP(hipStreamSynchronize(c->stream));
  graph_release(c);
  const bool timing = c->timing;
  c->timing = false;
  int rc = 0;
  for (int r = 0; r < 2 && rc == 0; ++r) {
    const int p = c->cur;
    rc = capture(c, [&] { return stage_a(c, budget); }, &c->graph_a[p]);
    if (rc) break;
    after_graph_a(c, p);
    rc = capture(c, [&] { return stage_b(c); }, &c->graph_b[1 - p]);
    positions_changed(c);
  }
  c->timing = timing;
  if (rc) {
    graph_release(c);
    return rc;
  }
  c->graph_budget = budget;
  c->graph_num = c->num;
  return 0;
}

// One two-phase step through the captured graphs (same result as
// sphx_mf_step).  Reports the iterations of the final solve and its error.
int sphx_graph_step(sphx_context* c, int* iters, double* err) {
  SPHX_CHECK(c, "sphx_graph_step", "null context");
  SPHX_CHECK(c->graph_num == c->num && c->graph_a[c->cur], "sphx_graph_step",
             "no step graph for this state (call sphx_graph_build)");
  SPHX_CHECK(iters || err || c->coop, "sphx_graph_step",
             "a step without outcome needs the persistent solve (sphx_persistent_solve)");
  SPHX_TRY(ensure_device(c));
  SPHX_TRY(mlp_async_settle(c));   // graph steps adopt a finished refresh at once
  const int p = c->cur;
  SPHX_HIP(hipGraphLaunch(c->graph_a[p], c->stream));
  after_graph_a(c, p);
  int it = 0;
  double e = 0.0;
  if (c->coop) {
    // the rest of the solve and the learned start's fallback on the device;
    // the host reads the outcome only when asked for it
    SPHX_TRY(enqueue_solve_coop(c, false));
    if (iters || err) {
      SPHX_TRY(read_solve(c, &it, &e));
      SPHX_TRY(mlp_fallback(c, &it, &e));
    }
  } else {
    SPHX_TRY(continue_solve(c, &it, &e));
    SPHX_TRY(mlp_fallback(c, &it, &e));
  }
  SPHX_HIP(hipGraphLaunch(c->graph_b[1 - p], c->stream));
  positions_changed(c);
  if (iters) *iters = it;
  if (err) *err = e;
  return 0;
}

// Persistent solve on or off (graphs are recaptured by the next
// sphx_graph_build).
int sphx_persistent_solve(sphx_context* c, int on) {
  SPHX_CHECK(c, "sphx_persistent_solve", "null context");
  SPHX_CHECK(!on || c->mlp.online_every == 0, "sphx_persistent_solve",
             "not available together with the learned start's online refresh");
  SPHX_TRY(ensure_device(c));
  if (c->coop != (on != 0)) graph_release(c);
  c->coop = on != 0;
  if (c->coop) {
    SPHX_TRY(solve_log_setup(c));
    SPHX_HIP(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int sphx_solve_history(sphx_context* c, int* iters, double* err, int max_solves, int* count) {
  SPHX_CHECK(c && count && max_solves >= 0, "sphx_solve_history", "bad argument");
  return solve_log_read(c, iters, err, max_solves, count);
}

int sphx_graph_release(sphx_context* c) {
  SPHX_CHECK(c, "sphx_graph_release", "null context");
  graph_release(c);
  return 0;
}

}  // extern "C"

namespace sphx {

static int tune_index(const char* name) {
  for (int t = 0; t < T_COUNT; ++t)
    if (std::strcmp(name, TUNE_NAMES[t]) == 0) return t;
  return -1;
}

// ---------------------------------------------------------- tuning cache --
// The file is a JSON array with one object per line,
//   {"arch": "gfx950", "bucket": 20, "kernel": "iisph_iter", "block": 128},
// written here and read back line by line (the reader accepts exactly what
// the writer produces).
struct TuneEntry {
  std::string arch, kernel;
  int bucket = 0, block = 0;
};

static int size_bucket(int n) {
  int b = 0;
  while (n > 1) { n >>= 1; ++b; }
  return b;
}

static int device_arch(sphx_context* c, std::string* arch) {
  hipDeviceProp_t prop;
  SPHX_HIP(hipGetDeviceProperties(&prop, c->device));
  std::string a = prop.gcnArchName;
  *arch = a.substr(0, a.find(':'));   // gfx950 (feature suffixes dropped)
  return 0;
}

static bool parse_str(const std::string& line, const char* key, std::string* out) {
  const std::string k = std::string("\"") + key + "\": \"";
  const size_t a = line.find(k);
  if (a == std::string::npos) return false;
  const size_t b = line.find('"', a + k.size());
  if (b == std::string::npos) return false;
  *out = line.substr(a + k.size(), b - a - k.size());
  return true;
}

static bool parse_int(const std::string& line, const char* key, int* out) {
  const std::string k = std::string("\"") + key + "\": ";
  const size_t a = line.find(k);
  if (a == std::string::npos) return false;
  char* end = nullptr;
  const long v = std::strtol(line.c_str() + a + k.size(), &end, 10);
  if (end == line.c_str() + a + k.size()) return false;
  *out = (int)v;
  return true;
}

static std::vector<TuneEntry> read_cache(const char* path) {
  std::vector<TuneEntry> out;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    TuneEntry e;
    if (parse_str(line, "arch", &e.arch) && parse_str(line, "kernel", &e.kernel) &&
        parse_int(line, "bucket", &e.bucket) && parse_int(line, "block", &e.block))
      out.push_back(e);
  }
  return out;
}

}  // namespace sphx

extern "C" {

int sphx_tune_set(sphx_context* c, const char* kernel, int block) {
  SPHX_CHECK(c && kernel, "sphx_tune_set", "null argument");
  int t = tune_index(kernel);
  SPHX_CHECK(t >= 0, "sphx_tune_set", "unknown kernel name");
  SPHX_CHECK(block >= 64 && block <= 256 && block % 64 == 0, "sphx_tune_set",
             "block must be 64, 128, 192 or 256");
  c->tune[t] = block;
  return 0;
}

int sphx_tune_get(sphx_context* c, const char* kernel, int* block) {
  SPHX_CHECK(c && kernel && block, "sphx_tune_get", "null argument");
  int t = tune_index(kernel);
  SPHX_CHECK(t >= 0, "sphx_tune_get", "unknown kernel name");
  *block = c->tune[t];
  return 0;
}

int sphx_tune_cache_save(sphx_context* c, const char* path) {
  SPHX_CHECK(c && path, "sphx_tune_cache_save", "null argument");
  std::string arch;
  SPHX_TRY(device_arch(c, &arch));
  const int bucket = size_bucket(c->num);
  std::vector<TuneEntry> all;
  for (const TuneEntry& e : read_cache(path))
    if (!(e.arch == arch && e.bucket == bucket && tune_index(e.kernel.c_str()) >= 0)) all.push_back(e);
  for (int t = 0; t < T_COUNT; ++t) all.push_back(TuneEntry{arch, TUNE_NAMES[t], bucket, c->tune[t]});
  std::ostringstream o;
  o << "[\n";
  for (size_t q = 0; q < all.size(); ++q)
    o << "  {\"arch\": \"" << all[q].arch << "\", \"bucket\": " << all[q].bucket
      << ", \"kernel\": \"" << all[q].kernel << "\", \"block\": " << all[q].block << "}"
      << (q + 1 < all.size() ? ",\n" : "\n");
  o << "]\n";
  std::ofstream f(path, std::ios::trunc);
  SPHX_CHECK(f.good(), "sphx_tune_cache_save", "cannot open the cache file for writing");
  f << o.str();
  f.close();
  SPHX_CHECK(f.good(), "sphx_tune_cache_save", "write failed");
  return 0;
}

int sphx_tune_cache_load(sphx_context* c, const char* path, int* applied) {
  SPHX_CHECK(c && path && applied, "sphx_tune_cache_load", "null argument");
  *applied = 0;
  std::string arch;
  SPHX_TRY(device_arch(c, &arch));
  const int bucket = size_bucket(c->num);
  for (const TuneEntry& e : read_cache(path)) {
    if (e.arch != arch || e.bucket != bucket) continue;
    const int t = tune_index(e.kernel.c_str());
    if (t < 0 || e.block < 64 || e.block > 256 || e.block % 64) continue;
    c->tune[t] = e.block;
    ++*applied;
  }
  if (*applied && c->graph_num >= 0) SPHX_TRY(sphx_graph_build(c, c->graph_budget));
  return 0;
}

// Project autotuner: times every neighbour-loop kernel class of the
// two-phase step at each candidate workgroup size over `steps` real steps
// (HIP events on the context stream) and keeps, per kernel, the size with
// the shortest mean launch.  The steps advance the simulation.  Step graphs
// are recaptured with the new sizes.  Block sizes do not change any
// particle's arithmetic.
int sphx_autotune(sphx_context* c, int steps) {
  SPHX_CHECK(c, "sphx_autotune", "null context");
  SPHX_CHECK(c->nphase > 0, "sphx_autotune", "tunes the two-phase step (call sphx_mf_setup)");
  SPHX_CHECK(steps >= 1, "sphx_autotune", "steps must be >= 1");
  const int cand[3] = {64, 128, 256};
  double best[T_COUNT];
  int pick[T_COUNT];
  for (int t = 0; t < T_COUNT; ++t) { best[t] = 1e300; pick[t] = c->tune[t]; }
  const bool timing = c->timing;
  const bool had_graph = c->graph_num >= 0;
  c->timing = true;
  for (int k = 0; k < 3; ++k) {
    for (int t = 0; t < T_COUNT; ++t) c->tune[t] = cand[k];
    SPHX_TRY(sphx_timing_reset(c));
    for (int s = 0; s < steps; ++s) SPHX_TRY(sphx_mf_step(c, nullptr, nullptr));
    SPHX_TRY(timer_flush(c));
    for (int t = 0; t < T_COUNT; ++t) {
      double ms = 0.0;
      int n = 0;
      SPHX_TRY(sphx_timing_get(c, TUNE_NAMES[t], &ms, &n));
      if (n > 0 && ms / n < best[t]) { best[t] = ms / n; pick[t] = cand[k]; }
    }
  }
  for (int t = 0; t < T_COUNT; ++t) c->tune[t] = pick[t];
  SPHX_TRY(sphx_timing_reset(c));
  c->timing = timing;
  if (had_graph) SPHX_TRY(sphx_graph_build(c, c->graph_budget));
  return 0;
}

}  // extern "C"
