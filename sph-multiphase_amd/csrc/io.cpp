// Host-side state I/O in the reference's formats:
//  - scene.xml parameters (FluidSystem::ParseXML, fluid_system.cpp:897-977,
//    with SetDruckerPragerParam :882-892 and the value syntax of
//    catpaw/catXMLHelper.cpp:7-57);
//  - the text frame (FluidSystem::outputFile, FluidIOUtility.cpp:17-38);
//  - checkpoints: the ParamCarrier and the displayPack / calculationPack
//    arrays exactly as laid out in common_header.h, so a state saved after
//    any step is reloaded through sphx_upload and continues identically.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sphx.h"

namespace {

thread_local std::string g_io_error;

}  // namespace

namespace sphx {
// Latch an I/O error message (read with sphx_io_last_error); returns -1.
int io_fail(const std::string& msg) {
  g_io_error = msg;
  return -1;
}
}  // namespace sphx

namespace {

using sphx::io_fail;

// The flat element structure scene.xml uses: a top-level element with a
// name and attributes, whose children are elements holding text.
struct Elem {
  std::string name;
  std::string attrs;   // raw attribute text of the opening tag
  size_t body_begin = 0, body_end = 0;
};

// Skip comments <!-- ... --> and declarations <? ... ?> starting at p.
size_t skip_misc(const std::string& s, size_t p) {
  for (;;) {
    size_t q = s.find('<', p);
    if (q == std::string::npos) return q;
    if (s.compare(q, 4, "<!--") == 0) {
      size_t e = s.find("-->", q + 4);
      if (e == std::string::npos) return std::string::npos;
      p = e + 3;
    } else if (s.compare(q, 2, "<?") == 0) {
      size_t e = s.find("?>", q + 2);
      if (e == std::string::npos) return std::string::npos;
      p = e + 2;
    } else {
      return q;
    }
  }
}

// Next element at or after p (its closing tag is the first </name> after the
// opening tag: scene.xml elements do not nest elements of their own name).
bool next_elem(const std::string& s, size_t& p, size_t end, Elem& e) {
  size_t q = skip_misc(s, p);
  if (q == std::string::npos || q >= end) return false;
  size_t gt = s.find('>', q);
  if (gt == std::string::npos || gt >= end) return false;
  std::string tag = s.substr(q + 1, gt - q - 1);
  bool self_closing = !tag.empty() && tag.back() == '/';
  if (self_closing) tag.pop_back();
  size_t sp = tag.find_first_of(" \t\r\n");
  e.name = tag.substr(0, sp);
  e.attrs = sp == std::string::npos ? "" : tag.substr(sp);
  if (self_closing) {
    e.body_begin = e.body_end = gt + 1;
    p = gt + 1;
    return true;
  }
  std::string close = "</" + e.name + ">";
  size_t c = s.find(close, gt + 1);
  if (c == std::string::npos || c >= end) return false;
  e.body_begin = gt + 1;
  e.body_end = c;
  p = c + close.size();
  return true;
}

// Children of a top-level element, by name (first occurrence, as
// tinyxml2's FirstChildElement).  Commented-out children are skipped.
bool child_text(const std::string& s, const Elem& parent, const char* name, std::string& out) {
  size_t p = parent.body_begin;
  Elem c;
  while (next_elem(s, p, parent.body_end, c)) {
    if (c.name == name) {
      out = s.substr(c.body_begin, c.body_end - c.body_begin);
      return true;
    }
  }
  return false;
}

bool attr_int(const std::string& attrs, const char* name, int* v) {
  std::string key = std::string(name) + "=\"";
  size_t p = attrs.find(key);
  if (p == std::string::npos) return false;
  return std::sscanf(attrs.c_str() + p + key.size(), "%d", v) == 1;
}

// XMLGetFloat: the child's text as a float, 0 when absent (catXMLHelper.cpp:17-25).
float get_float(const std::string& s, const Elem* e, const char* name) {
  std::string t;
  float v = 0.f;
  if (e && child_text(s, *e, name, t)) std::sscanf(t.c_str(), "%f", &v);
  return v;
}

int get_int(const std::string& s, const Elem* e, const char* name) {
  std::string t;
  int v = 0;
  if (e && child_text(s, *e, name, t)) std::sscanf(t.c_str(), "%d", &v);
  return v;
}

// QueryFloat3: "x,y,z" (catXMLHelper.cpp:7-14), zero when absent.
sphx_float3 get_float3(const std::string& s, const Elem* e, const char* name) {
  std::string t;
  sphx_float3 v = {0.f, 0.f, 0.f};
  if (e && child_text(s, *e, name, t)) {
    const char* c = t.c_str();
    while (*c == ' ' || *c == '\t' || *c == '\r' || *c == '\n') ++c;
    std::sscanf(c, "%f,%f,%f", &v.x, &v.y, &v.z);
  }
  return v;
}

// XMLGetFloatN with N = 3: "a,b,c" (catXMLHelper.cpp:46-57).
void get_float_n3(const std::string& s, const Elem* e, const char* name, float* out) {
  std::string t;
  if (e && child_text(s, *e, name, t)) {
    const char* c = t.c_str();
    while (*c == ' ' || *c == '\t' || *c == '\r' || *c == '\n') ++c;
    std::sscanf(c, "%f,%f,%f", &out[0], &out[1], &out[2]);
  }
}

constexpr char kCheckpointMagic[8] = {'S', 'P', 'H', 'X', 'C', 'K', 'P', '1'};


}  // namespace

extern "C" {

const char* sphx_io_last_error(void) { return g_io_error.c_str(); }

int sphx_scene_xml_parse(const char* text, int caseid, sphx_ParamCarrier* P,
                         sphx_SceneXml* scene) {
  if (!text || !P) return io_fail("sphx_scene_xml_parse: null argument");
  const std::string s(text);
  std::vector<Elem> top;
  {
    size_t p = 0;
    Elem e;
    while (next_elem(s, p, s.size(), e)) top.push_back(e);
  }
  const Elem* fluid = nullptr;
  const Elem* sc = nullptr;
  const Elem* bound = nullptr;
  for (const Elem& e : top) {
    if (e.name == "Fluid" && !fluid) fluid = &e;
    if (e.name == "BoundInfo" && !bound) bound = &e;
  }
  if (!fluid) return io_fail("sphx_scene_xml_parse: missing Fluid node");
  bool any_scene = false;
  for (const Elem& e : top) {
    if (e.name != "MultiScene") continue;
    any_scene = true;
    int id = 0;
    if (attr_int(e.attrs, "id", &id) && id == caseid) {
      sc = &e;
      break;
    }
  }
  if (!any_scene) return io_fail("sphx_scene_xml_parse: missing MultiScene node");
  if (!sc) return io_fail("sphx_scene_xml_parse: scene " + std::to_string(caseid) + " not found");

  // fluid_system.cpp:931-936
  P->acclimit = get_float(s, fluid, "AccelLimit");
  P->vlimit = get_float(s, fluid, "VelLimit");
  P->gravity = get_float3(s, fluid, "Gravity");
  P->volmin = get_float3(s, fluid, "VolMin");
  P->volmax = get_float3(s, fluid, "VolMax");

  // :939-968
  sphx_SceneXml tmp;
  std::memset(&tmp, 0, sizeof(tmp));
  sphx_SceneXml& X = scene ? *scene : tmp;
  X.maxPointNum = get_int(s, sc, "Num");
  X.scaleP = get_float(s, sc, "ScaleP");
  X.fluidnum = get_int(s, sc, "FluidCount");
  X.fvmin[0] = get_float3(s, sc, "VolMin0");
  X.fvmax[0] = get_float3(s, sc, "VolMax0");
  X.fvmin[1] = get_float3(s, sc, "VolMin1");
  X.fvmax[1] = get_float3(s, sc, "VolMax1");
  get_float_n3(s, sc, "MassRatio", X.massratio);
  get_float_n3(s, sc, "DensityRatio", X.densratio);
  get_float_n3(s, sc, "ViscRatio", X.viscratio);

  P->maxNum = X.maxPointNum;
  P->simscale = get_float(s, sc, "SimScale");
  P->dt = get_float(s, sc, "DT");
  P->smoothradius = get_float(s, sc, "SmoothRadius");
  P->viscosity = get_float(s, sc, "Viscosity");
  P->mass = get_float(s, sc, "Mass");
  P->restdensity = get_float(s, sc, "RestDensity");
  P->radius = get_float(s, sc, "Radius");
  P->softminx = get_float3(s, sc, "SoftBoundMin");
  P->softmaxx = get_float3(s, sc, "SoftBoundMax");
  P->solidK = get_float(s, sc, "SolidK");
  P->solidG = get_float(s, sc, "SolidG");
  P->boundstiff = get_float(s, sc, "BoundStiff");
  // Drucker-Prager constants from the angles in degrees (:882-892)
  float phi = get_float(s, sc, "DP_Phi"), psi = get_float(s, sc, "DP_Psi");
  float kc = get_float(s, sc, "DP_cohesion");
  phi *= 3.141593f / 180.0f;
  psi *= 3.141593f / 180.0f;
  P->a_phi = (float)(2 * std::sin(phi) / std::sqrt(3.0) / (3 - std::sin(phi)));
  P->a_psi = (float)(2 * std::sin(psi) / std::sqrt(3.0) / (3 - std::sin(psi)));
  P->k_c = (float)(3 * kc * std::cos(psi) / std::sqrt(3.0) / (3 - std::sin(phi)));

  // :972-975
  P->bmass = get_float(s, bound, "Mass");
  P->bRestdensity = get_float(s, bound, "RestDensity");
  P->bvisc = get_float(s, bound, "Viscosity");
  return 0;
}

int sphx_scene_xml_load(const char* path, int caseid, sphx_ParamCarrier* P,
                        sphx_SceneXml* scene) {
  if (!path) return io_fail("sphx_scene_xml_load: null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(std::string("sphx_scene_xml_load: cannot open ") + path);
  std::string text;
  char buf[65536];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  std::fclose(f);
  return sphx_scene_xml_parse(text.c_str(), caseid, P, scene);
}

// outputFile (FluidIOUtility.cpp:17-38): the count, then per particle
// "x y z mass restdens type" with %f / %d.
int sphx_frame_write(const char* path, const sphx_displayPack* disp,
                     const sphx_calculationPack* calc, int n) {
  if (!path || n < 0 || (n > 0 && (!disp || !calc)))
    return io_fail("sphx_frame_write: bad argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return io_fail(std::string("sphx_frame_write: cannot open ") + path);
  std::fprintf(f, "%d\n", n);
  for (int i = 0; i < n; ++i) {
    std::fprintf(f, "%f %f %f ", disp[i].pos.x, disp[i].pos.y, disp[i].pos.z);
    std::fprintf(f, "%f %f %d\n", calc[i].mass, calc[i].restdens, disp[i].type);
  }
  if (std::fclose(f) != 0) return io_fail("sphx_frame_write: write failed");
  return 0;
}

// Read a frame back: fills pos, mass, restdens and type of the first
// min(count, cap) particles; *n receives the frame's count.
int sphx_frame_read(const char* path, sphx_displayPack* disp, sphx_calculationPack* calc,
                    int cap, int* n) {
  if (!path || cap < 0) return io_fail("sphx_frame_read: bad argument");
  FILE* f = std::fopen(path, "r");
  if (!f) return io_fail(std::string("sphx_frame_read: cannot open ") + path);
  int count = 0;
  if (std::fscanf(f, "%d", &count) != 1 || count < 0) {
    std::fclose(f);
    return io_fail("sphx_frame_read: bad header");
  }
  if (n) *n = count;
  const int m = count < cap ? count : cap;
  for (int i = 0; i < m; ++i) {
    float x, y, z, mass, rest;
    int type;
    if (std::fscanf(f, "%f %f %f %f %f %d", &x, &y, &z, &mass, &rest, &type) != 6) {
      std::fclose(f);
      return io_fail("sphx_frame_read: truncated at particle " + std::to_string(i));
    }
    if (disp) {
      disp[i].pos.x = x; disp[i].pos.y = y; disp[i].pos.z = z;
      disp[i].type = type;
    }
    if (calc) {
      calc[i].mass = mass;
      calc[i].restdens = rest;
    }
  }
  std::fclose(f);
  return 0;
}

// Checkpoint: magic, the three struct sizes, n, the ParamCarrier, then the
// displayPack and calculationPack arrays, all in the reference's layout.
int sphx_checkpoint_write(const char* path, const sphx_ParamCarrier* P,
                          const sphx_displayPack* disp, const sphx_calculationPack* calc,
                          int n) {
  if (!path || !P || n < 0 || (n > 0 && (!disp || !calc)))
    return io_fail("sphx_checkpoint_write: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return io_fail(std::string("sphx_checkpoint_write: cannot open ") + path);
  const int32_t head[4] = {(int32_t)sizeof(sphx_ParamCarrier), (int32_t)sizeof(sphx_displayPack),
                           (int32_t)sizeof(sphx_calculationPack), (int32_t)n};
  bool ok = std::fwrite(kCheckpointMagic, 1, 8, f) == 8 &&
            std::fwrite(head, sizeof(head), 1, f) == 1 &&
            std::fwrite(P, sizeof(*P), 1, f) == 1 &&
            (n == 0 || (std::fwrite(disp, sizeof(*disp), (size_t)n, f) == (size_t)n &&
                        std::fwrite(calc, sizeof(*calc), (size_t)n, f) == (size_t)n));
  if (std::fclose(f) != 0) ok = false;
  return ok ? 0 : io_fail("sphx_checkpoint_write: write failed");
}

// Reads a checkpoint: *n receives its particle count; the particles are
// copied when cap >= n (call with cap = 0 to size the arrays first).
int sphx_checkpoint_read(const char* path, sphx_ParamCarrier* P, sphx_displayPack* disp,
                         sphx_calculationPack* calc, int cap, int* n) {
  if (!path) return io_fail("sphx_checkpoint_read: null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(std::string("sphx_checkpoint_read: cannot open ") + path);
  char magic[8];
  int32_t head[4];
  sphx_ParamCarrier p;
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kCheckpointMagic, 8) != 0 ||
      std::fread(head, sizeof(head), 1, f) != 1) {
    std::fclose(f);
    return io_fail("sphx_checkpoint_read: not a checkpoint");
  }
  if (head[0] != (int32_t)sizeof(sphx_ParamCarrier) ||
      head[1] != (int32_t)sizeof(sphx_displayPack) ||
      head[2] != (int32_t)sizeof(sphx_calculationPack) || head[3] < 0) {
    std::fclose(f);
    return io_fail("sphx_checkpoint_read: layout differs from this build's");
  }
  if (std::fread(&p, sizeof(p), 1, f) != 1) {
    std::fclose(f);
    return io_fail("sphx_checkpoint_read: truncated parameters");
  }
  const int count = head[3];
  if (n) *n = count;
  if (P) *P = p;
  if (cap >= count && count > 0 && disp && calc) {
    if (std::fread(disp, sizeof(*disp), (size_t)count, f) != (size_t)count ||
        std::fread(calc, sizeof(*calc), (size_t)count, f) != (size_t)count) {
      std::fclose(f);
      return io_fail("sphx_checkpoint_read: truncated particles");
    }
  }
  std::fclose(f);
  return 0;
}

// Boundary particles from a shape file (FluidSystem::LoadBoundary,
// AddShape.cpp:265-296): "WALL min max" and "OPEN_BOX min max thickness"
// commands ("x,y,z" vectors), each wall filled as SetupAddWall does
// (:298-346): spacing (bmass / bRestdensity)^(1/3) / simscale
// (CalcBoundarySpacing :257-263), ceil(extent / spacing) columns in x and
// z from min, layers y = min.y, min.y + spacing, ... <= max.y (a float
// accumulator, as there).  Particles are appended from slot `start` up to
// `cap` (AddParticle's defaults, :702-717, with bornid = slot); *added
// receives the number the file asks for, also beyond cap.
int sphx_boundary_parse(const char* text, const sphx_ParamCarrier* P, sphx_displayPack* disp,
                        sphx_calculationPack* calc, int start, int cap, int* added) {
  if (!text || !P || start < 0 || cap < start) return io_fail("sphx_boundary_parse: bad argument");
  const float spacing = std::pow(P->bmass / P->bRestdensity, 1 / 3.0f) / P->simscale;
  if (!(spacing > 0.f) || !std::isfinite(spacing))
    return io_fail("sphx_boundary_parse: bmass / bRestdensity / simscale give no spacing");
  // more particles than any buffer could hold is a malformed file
  constexpr double kMaxParticles = double(1 << 30);
  int n = start;
  const char* bad = nullptr;
  auto add_wall = [&](sphx_float3 mn, sphx_float3 mx) {
    if (bad) return;
    const float ex = std::ceil((mx.x - mn.x) / spacing), ez = std::ceil((mx.z - mn.z) / spacing);
    const double ey = std::floor(((double)mx.y - (double)mn.y) / spacing) + 1.0;
    if (!std::isfinite(ex) || !std::isfinite(ez) || !std::isfinite(ey) || !std::isfinite(mn.y)) {
      bad = "sphx_boundary_parse: wall bounds are not finite";
      return;
    }
    const int cntx = ex > 0.f ? (int)std::fmin(ex, (float)kMaxParticles) : 0;
    const int cntz = ez > 0.f ? (int)std::fmin(ez, (float)kMaxParticles) : 0;
    if ((double)n + (double)cntx * cntz * std::fmax(ey, 0.0) > kMaxParticles) {
      bad = "sphx_boundary_parse: the walls describe more than 2^30 particles";
      return;
    }
    const int cnt = cntx * cntz;
    for (float y = mn.y; y <= mx.y; y += spacing) {
      if (y + spacing == y) {
        bad = "sphx_boundary_parse: wall too far from the origin for its spacing";
        return;
      }
      for (int xz = 0; xz < cnt; ++xz) {
        const float x = mn.x + (xz % cntx) * spacing;
        const float z = mn.z + (xz / cntx) * spacing;
        if (n < cap) {
          if (disp) {
            sphx_displayPack& d = disp[n];
            std::memset(&d, 0, sizeof(d));
            d.pos.x = x; d.pos.y = y; d.pos.z = z;
            d.color.x = 1.f; d.color.y = 1.f; d.color.z = 1.f;
            d.color.w = z > 0 ? 0.f : 0.2f;
            d.type = SPHX_TYPE_BOUNDARY;
          }
          if (calc) {
            sphx_calculationPack& c = calc[n];
            std::memset(&c, 0, sizeof(c));
            c.bornid = n;
            c.deformGrad.data[0] = c.deformGrad.data[4] = c.deformGrad.data[8] = 1.f;
            c.mass = P->bmass;
            c.restdens = P->bRestdensity;
            c.visc = P->bvisc;
          }
        }
        ++n;
      }
    }
  };
  auto read3 = [](const char*& c, sphx_float3& v) {
    int used = 0;
    if (std::sscanf(c, " %f , %f , %f%n", &v.x, &v.y, &v.z, &used) != 3) return false;
    c += used;
    return true;
  };
  const char* c = text;
  char word[1000];
  int used = 0;
  while (std::sscanf(c, " %999s%n", word, &used) == 1) {
    c += used;
    sphx_float3 mn, mx;
    if (!std::strcmp(word, "WALL")) {
      if (!read3(c, mn) || !read3(c, mx)) return io_fail("sphx_boundary_parse: bad WALL");
      add_wall(mn, mx);
    } else if (!std::strcmp(word, "OPEN_BOX")) {
      float t = 0.f;
      if (!read3(c, mn) || !read3(c, mx) || std::sscanf(c, " %f%n", &t, &used) != 1)
        return io_fail("sphx_boundary_parse: bad OPEN_BOX");
      c += used;
      // SetupAddOpenBox (:348-359): ground, front, back, left, right
      add_wall({mn.x, mn.y - t, mn.z}, {mx.x, mn.y, mx.z});
      add_wall({mn.x - t, mn.y, mn.z - t}, {mx.x + t, mx.y, mn.z});
      add_wall({mn.x - t, mn.y, mx.z}, {mx.x + t, mx.y, mx.z + t});
      add_wall({mn.x - t, mn.y, mn.z}, {mn.x, mx.y, mx.z});
      add_wall({mx.x, mn.y, mn.z}, {mx.x + t, mx.y, mx.z});
    }
    if (bad) return io_fail(bad);
  }
  if (added) *added = n - start;
  return 0;
}

int sphx_boundary_load(const char* path, const sphx_ParamCarrier* P, sphx_displayPack* disp,
                       sphx_calculationPack* calc, int start, int cap, int* added) {
  if (!path) return io_fail("sphx_boundary_load: null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(std::string("sphx_boundary_load: cannot open ") + path);
  std::string text;
  char buf[65536];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  std::fclose(f);
  return sphx_boundary_parse(text.c_str(), P, disp, calc, start, cap, added);
}

}  // extern "C"
