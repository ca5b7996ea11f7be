// Checkpoints of a context (the parts of the state I/O that need the
// device): sphx_checkpoint_save writes the reference's AoS checkpoint
// (sphx_checkpoint_write, io.cpp) and appends tagged sections for the state
// the AoS arrays do not hold; sphx_checkpoint_restore reads both back.
// Section layout after the particles: 8-byte tag, int64 byte count,
// payload; unknown tags are skipped.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sphx.h"

namespace sphx {
int io_fail(const std::string& msg);
// context state beyond the AoS arrays (multiphase.hip, mlp.hip)
int mf_state_get(sphx_context* c, std::vector<char>& out);
int mf_state_set(sphx_context* c, const char* p, size_t bytes);
int mlp_state_get(sphx_context* c, std::vector<char>& out);
int mlp_state_set(sphx_context* c, const char* p, size_t bytes);
}  // namespace sphx

namespace {

using sphx::io_fail;

constexpr char kSectionMF[8] = {'S', 'P', 'H', 'X', 'M', 'F', '0', '1'};
constexpr char kSectionMLP[8] = {'S', 'P', 'H', 'X', 'N', 'E', 'T', '1'};

}  // namespace

extern "C" {

// A context's state to a checkpoint: its parameter block and every
// particle in slot order (stable cell-sorted after a step).
int sphx_checkpoint_save(sphx_context* ctx, const char* path) {
  if (!ctx || !path) return io_fail("sphx_checkpoint_save: null argument");
  sphx_ParamCarrier P;
  if (sphx_get_params(ctx, &P) != 0) return io_fail(sphx_last_error());
  const int n = sphx_num_particles(ctx);
  if (n < 0) return io_fail(sphx_last_error());
  std::vector<sphx_displayPack> disp((size_t)n);
  std::vector<sphx_calculationPack> calc((size_t)n);
  if (n > 0 && sphx_download(ctx, disp.data(), calc.data(), 0, n) != 0)
    return io_fail(sphx_last_error());
  P.num = n;
  // state the AoS arrays do not hold: multiphase fractions and g - a, the
  // network (weights, Adam moments, per-particle history, solve counter)
  std::vector<char> mf, net;
  if (sphx::mf_state_get(ctx, mf) != 0 || sphx::mlp_state_get(ctx, net) != 0)
    return io_fail(sphx_last_error());
  if (sphx_checkpoint_write(path, &P, disp.data(), calc.data(), n) != 0) return -1;
  if (mf.empty() && net.empty()) return 0;
  FILE* f = std::fopen(path, "ab");
  if (!f) return io_fail(std::string("sphx_checkpoint_save: cannot reopen ") + path);
  bool ok = true;
  auto section = [&](const char* tag, const std::vector<char>& b) {
    if (b.empty()) return;
    const int64_t len = (int64_t)b.size();
    ok = ok && std::fwrite(tag, 1, 8, f) == 8 && std::fwrite(&len, sizeof(len), 1, f) == 1 &&
         std::fwrite(b.data(), 1, b.size(), f) == b.size();
  };
  section(kSectionMF, mf);
  section(kSectionMLP, net);
  if (std::fclose(f) != 0) ok = false;
  return ok ? 0 : io_fail("sphx_checkpoint_save: write failed");
}

// Resume from a checkpoint into a context created for at least its
// particle count: parameters, then the particles (which become slots
// 0..n-1, num = n).
int sphx_checkpoint_restore(sphx_context* ctx, const char* path) {
  if (!ctx || !path) return io_fail("sphx_checkpoint_restore: null argument");
  sphx_ParamCarrier P;
  int n = 0;
  if (sphx_checkpoint_read(path, &P, nullptr, nullptr, 0, &n) != 0) return -1;
  sphx_ParamCarrier cur;
  if (sphx_get_params(ctx, &cur) != 0) return io_fail(sphx_last_error());
  if (n > cur.maxNum)
    return io_fail("sphx_checkpoint_restore: checkpoint holds more particles than the context");
  std::vector<sphx_displayPack> disp((size_t)n);
  std::vector<sphx_calculationPack> calc((size_t)n);
  if (sphx_checkpoint_read(path, &P, disp.data(), calc.data(), n, &n) != 0) return -1;
  P.maxNum = cur.maxNum;
  P.num = 0;
  if (sphx_set_params(ctx, &P) != 0) return io_fail(sphx_last_error());
  if (n > 0 && sphx_upload(ctx, disp.data(), calc.data(), 0, n, 1) != 0)
    return io_fail(sphx_last_error());
  // tagged sections after the particles (unknown tags are skipped)
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(std::string("sphx_checkpoint_restore: cannot reopen ") + path);
  const long skip = 8 + 4 * (long)sizeof(int32_t) + (long)sizeof(sphx_ParamCarrier) +
                    (long)n * (long)(sizeof(sphx_displayPack) + sizeof(sphx_calculationPack));
  int rc = std::fseek(f, skip, SEEK_SET) == 0 ? 0 : -1;
  char tag[8];
  while (rc == 0 && std::fread(tag, 1, 8, f) == 8) {
    int64_t len = 0;
    if (std::fread(&len, sizeof(len), 1, f) != 1 || len < 0) {
      rc = io_fail("sphx_checkpoint_restore: truncated section header");
      break;
    }
    std::vector<char> b((size_t)len);
    if (len > 0 && std::fread(b.data(), 1, (size_t)len, f) != (size_t)len) {
      rc = io_fail("sphx_checkpoint_restore: truncated section");
      break;
    }
    if (std::memcmp(tag, kSectionMF, 8) == 0) {
      if (sphx::mf_state_set(ctx, b.data(), b.size()) != 0) rc = io_fail(sphx_last_error());
    } else if (std::memcmp(tag, kSectionMLP, 8) == 0) {
      if (sphx::mlp_state_set(ctx, b.data(), b.size()) != 0) rc = io_fail(sphx_last_error());
    }
  }
  std::fclose(f);
  return rc;
}

}  // extern "C"
