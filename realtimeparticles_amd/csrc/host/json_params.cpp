// The JSON parameter channel of the reference's models: the trees of initFluidsJson
// (physics/ocl/Fluids.cpp:52-74) and initBoidsJson (physics/ocl/Boids.cpp:41-73), read the
// way transferJsonInputsToModel reads them (Fluids.cpp:218-251, Boids.cpp:177-210) and
// written back in the same shape for a parameter widget.
//
// Every numeric entry is an array [value, min, max] of which the value is used.  Value
// conversions follow the reference's JSON library (nlohmann_json 3.11.3, conanfile.py:12):
// a number read as float or integer is cast (truncating to an integer), a boolean read as a
// number is 0 or 1, an "Enable##..." switch must be a boolean (the boids read it as a bool;
// the fluids compare it with true, so a non-boolean reads as off there), and a missing key
// or a non-array where an array is read is an error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "rtp.h"

namespace rtp {
int fail(int status, const std::string& msg);  // sets the text rtp_last_error returns
}

namespace {

thread_local std::string g_jsonError;  // the message of the failure being reported

int bad() { return rtp::fail(RTP_ERR_INVALID, g_jsonError); }

struct JVal {
  enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
  bool b = false;
  double n = 0.0;
  std::string s;
  std::vector<JVal> a;
  std::vector<std::pair<std::string, JVal>> o;

  const JVal* get(const char* key) const
  {
    if (type != Obj)
      return nullptr;
    for (const auto& kv : o)
      if (kv.first == key)
        return &kv.second;
    return nullptr;
  }
};

// Recursive-descent reader of RFC 8259 JSON (UTF-8 passed through, \u escapes of the BMP).
struct Reader {
  const char* p;
  const char* end;
  std::string err;

  void ws()
  {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r'))
      ++p;
  }
  bool fail(const char* m)
  {
    if (err.empty())
      err = m;
    return false;
  }
  bool lit(const char* w)
  {
    const size_t l = std::strlen(w);
    if ((size_t)(end - p) < l || std::strncmp(p, w, l) != 0)
      return fail("invalid literal");
    p += l;
    return true;
  }
  bool str(std::string& out)
  {
    if (p >= end || *p != '"')
      return fail("expected a string");
    ++p;
    while (p < end && *p != '"') {
      char c = *p++;
      if (c != '\\') {
        out.push_back(c);
        continue;
      }
      if (p >= end)
        return fail("bad escape");
      c = *p++;
      switch (c) {
      case '"': case '\\': case '/': out.push_back(c); break;
      case 'b': out.push_back('\b'); break;
      case 'f': out.push_back('\f'); break;
      case 'n': out.push_back('\n'); break;
      case 'r': out.push_back('\r'); break;
      case 't': out.push_back('\t'); break;
      case 'u': {
        if (end - p < 4)
          return fail("bad \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
          const char h = *p++;
          v <<= 4;
          if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
          else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
          else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
          else return fail("bad \\u escape");
        }
        if (v < 0x80) {
          out.push_back((char)v);
        } else if (v < 0x800) {
          out.push_back((char)(0xc0 | (v >> 6)));
          out.push_back((char)(0x80 | (v & 0x3f)));
        } else {
          out.push_back((char)(0xe0 | (v >> 12)));
          out.push_back((char)(0x80 | ((v >> 6) & 0x3f)));
          out.push_back((char)(0x80 | (v & 0x3f)));
        }
        break;
      }
      default: return fail("bad escape");
      }
    }
    if (p >= end)
      return fail("unterminated string");
    ++p;
    return true;
  }
  bool value(JVal& v, int depth)
  {
    if (depth > 64)
      return fail("nesting too deep");
    ws();
    if (p >= end)
      return fail("unexpected end");
    const char c = *p;
    if (c == '{') {
      ++p;
      v.type = JVal::Obj;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      for (;;) {
        ws();
        std::string k;
        if (!str(k))
          return false;
        ws();
        if (p >= end || *p != ':')
          return fail("expected ':'");
        ++p;
        JVal x;
        if (!value(x, depth + 1))
          return false;
        v.o.emplace_back(std::move(k), std::move(x));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p;
      v.type = JVal::Arr;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      for (;;) {
        JVal x;
        if (!value(x, depth + 1))
          return false;
        v.a.push_back(std::move(x));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.type = JVal::Str;
      return str(v.s);
    }
    if (c == 't') {
      v.type = JVal::Bool;
      v.b = true;
      return lit("true");
    }
    if (c == 'f') {
      v.type = JVal::Bool;
      v.b = false;
      return lit("false");
    }
    if (c == 'n') {
      v.type = JVal::Null;
      return lit("null");
    }
    if (c == '-' || (c >= '0' && c <= '9')) {
      std::string num;
      while (p < end && (std::strchr("+-0123456789.eE", *p) != nullptr))
        num.push_back(*p++);
      char* stop = nullptr;
      v.type = JVal::Num;
      v.n = std::strtod(num.c_str(), &stop);
      if (stop == nullptr || *stop != '\0')
        return fail("bad number");
      return true;
    }
    return fail("unexpected character");
  }
};

bool parse(const char* text, JVal& root)
{
  if (!text) {
    g_jsonError = "NULL JSON text";
    return false;
  }
  Reader r{ text, text + std::strlen(text), {} };
  if (!r.value(root, 0)) {
    g_jsonError = "malformed JSON: " + r.err;
    return false;
  }
  r.ws();
  if (r.p != r.end) {
    g_jsonError = "malformed JSON: trailing characters";
    return false;
  }
  return true;
}

// A path lookup that records which key was missing.
struct Path {
  const JVal* v;
  std::string where;
  Path at(const char* key) const
  {
    const JVal* x = v ? v->get(key) : nullptr;
    return Path{ x, where + "/" + key };
  }
};

bool missing(const Path& q)
{
  if (q.v)
    return false;
  g_jsonError = "missing key " + q.where;
  return true;
}

// json[...][0] read as a number (a boolean reads as 0 or 1).
bool first_number(const Path& q, double* out)
{
  if (missing(q))
    return false;
  if (q.v->type != JVal::Arr || q.v->a.empty()) {
    g_jsonError = q.where + " is not a [value, min, max] array";
    return false;
  }
  const JVal& x = q.v->a[0];
  if (x.type == JVal::Num)
    *out = x.n;
  else if (x.type == JVal::Bool)
    *out = x.b ? 1.0 : 0.0;
  else {
    g_jsonError = q.where + "[0] is not a number";
    return false;
  }
  return true;
}

// A switch read as a bool (Boids.cpp:189-191, :194).
bool as_bool(const Path& q, bool* out)
{
  if (missing(q))
    return false;
  if (q.v->type != JVal::Bool) {
    g_jsonError = q.where + " is not a boolean";
    return false;
  }
  *out = q.v->b;
  return true;
}

// A switch compared with true (Fluids.cpp:237, :242): anything but the boolean true is off.
bool equals_true(const Path& q, bool* out)
{
  if (missing(q))
    return false;
  *out = q.v->type == JVal::Bool && q.v->b;
  return true;
}

// A value read as a number (Boids.cpp:202 reads "Attract" into an int).
bool as_number(const Path& q, double* out)
{
  if (missing(q))
    return false;
  if (q.v->type == JVal::Num)
    *out = q.v->n;
  else if (q.v->type == JVal::Bool)
    *out = q.v->b ? 1.0 : 0.0;
  else {
    g_jsonError = q.where + " is not a number";
    return false;
  }
  return true;
}

int32_t to_int(double v)
{
  if (!(v > -2147483648.0))
    return INT32_MIN;
  if (!(v < 2147483647.0))
    return INT32_MAX;
  return (int32_t)v;
}

int put(char* buf, size_t cap, size_t* len, const std::string& s)
{
  if (len)
    *len = s.size();
  if (buf && cap > 0) {
    const size_t k = s.size() < cap - 1 ? s.size() : cap - 1;
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
    if (k < s.size()) {
      g_jsonError = "buffer too small for the JSON text";
      return bad();
    }
  }
  return RTP_OK;
}

std::string num(double v)
{
  char t[40];
  std::snprintf(t, sizeof(t), "%.9g", v);
  return t;
}

std::string range3(double v, double lo, double hi) { return "[" + num(v) + ", " + num(lo) + ", " + num(hi) + "]"; }

// ---- the tree operations of the reference's Model (physics/Model.hpp:154-172), on JVal

// Value equality as the reference's JSON library defines it: objects compare as maps (key
// order does not matter; a repeated key counts with its last value, as the library keeps
// it), arrays element-wise, numbers by value.
const JVal* last_of(const JVal& obj, const std::string& key)
{
  const JVal* r = nullptr;
  for (const auto& kv : obj.o)
    if (kv.first == key)
      r = &kv.second;
  return r;
}

bool same(const JVal& a, const JVal& b)
{
  if (a.type != b.type)
    return false;
  switch (a.type) {
  case JVal::Null: return true;
  case JVal::Bool: return a.b == b.b;
  case JVal::Num: return a.n == b.n;
  case JVal::Str: return a.s == b.s;
  case JVal::Arr:
    if (a.a.size() != b.a.size())
      return false;
    for (size_t k = 0; k < a.a.size(); ++k)
      if (!same(a.a[k], b.a[k]))
        return false;
    return true;
  case JVal::Obj: {
    size_t na = 0, nb = 0;
    for (size_t k = 0; k < a.o.size(); ++k) {
      if (last_of(a, a.o[k].first) != &a.o[k].second)
        continue;  // an earlier duplicate
      ++na;
      const JVal* x = last_of(b, a.o[k].first);
      if (!x || !same(a.o[k].second, *x))
        return false;
    }
    for (size_t k = 0; k < b.o.size(); ++k)
      if (last_of(b, b.o[k].first) == &b.o[k].second)
        ++nb;
    return na == nb;
  }
  }
  return false;
}

// RFC 7386 JSON merge patch, the library's merge_patch: an object patch merges key by key
// (null removes the key), anything else replaces the target.
void merge_patch(JVal& target, const JVal& patch)
{
  if (patch.type != JVal::Obj) {
    target = patch;
    return;
  }
  if (target.type != JVal::Obj) {
    target = JVal{};
    target.type = JVal::Obj;
  }
  for (const auto& kv : patch.o) {
    auto it = target.o.begin();
    while (it != target.o.end() && it->first != kv.first)
      ++it;
    if (kv.second.type == JVal::Null) {
      // remove every entry of the key
      for (size_t k = 0; k < target.o.size();)
        if (target.o[k].first == kv.first)
          target.o.erase(target.o.begin() + (long)k);
        else
          ++k;
      continue;
    }
    if (it == target.o.end()) {
      target.o.emplace_back(kv.first, JVal{});
      merge_patch(target.o.back().second, kv.second);
    } else {
      merge_patch(it->second, kv.second);
    }
  }
}

void write_string(std::string& out, const std::string& s)
{
  out.push_back('"');
  for (const char ch : s) {
    const unsigned char c = (unsigned char)ch;
    if (c == '"' || c == '\\') {
      out.push_back('\\');
      out.push_back(ch);
    } else if (c < 0x20) {
      char t[8];
      std::snprintf(t, sizeof(t), "\\u%04x", c);
      out += t;
    } else {
      out.push_back(ch);
    }
  }
  out.push_back('"');
}

// Numbers are written with 17 significant digits, so every double reads back exactly.
void write(std::string& out, const JVal& v)
{
  switch (v.type) {
  case JVal::Null: out += "null"; break;
  case JVal::Bool: out += v.b ? "true" : "false"; break;
  case JVal::Num: {
    char t[40];
    if (std::isfinite(v.n) && v.n == std::floor(v.n) && std::fabs(v.n) < 1e15)
      std::snprintf(t, sizeof(t), "%.0f", v.n);
    else
      std::snprintf(t, sizeof(t), "%.17g", v.n);
    out += t;
    break;
  }
  case JVal::Str: write_string(out, v.s); break;
  case JVal::Arr:
    out.push_back('[');
    for (size_t k = 0; k < v.a.size(); ++k) {
      if (k)
        out += ", ";
      write(out, v.a[k]);
    }
    out.push_back(']');
    break;
  case JVal::Obj:
    out.push_back('{');
    for (size_t k = 0; k < v.o.size(); ++k) {
      if (k)
        out += ", ";
      write_string(out, v.o[k].first);
      out += ": ";
      write(out, v.o[k].second);
    }
    out.push_back('}');
    break;
  }
}

// The "Fluids" tree of initFluidsJson (Fluids.cpp:52-74) and of initCloudsJson
// (Clouds.cpp:65-87): FluidKernelInputs and the iteration count (Fluids.cpp:229-246).
bool read_fluids(const JVal& root, rtp_fluid_params* p, uint32_t* iters)
{
  const Path f = Path{ &root, "" }.at("Fluids");
  const Path ap = f.at("Artificial Pressure"), vc = f.at("Vorticity Confinement");
  double it, rho, cfm, dt, apc, apr, ape, vcc, xs;
  bool apOn, vcOn;
  if (!first_number(f.at("Nb Jacobi Iterations"), &it) || !first_number(f.at("Rest Density"), &rho) ||
      !first_number(f.at("Relax CFM"), &cfm) || !first_number(f.at("Time Step"), &dt) ||
      !equals_true(ap.at("Enable##Pressure"), &apOn) ||
      !first_number(ap.at("Coefficient##Pressure"), &apc) || !first_number(ap.at("Radius"), &apr) ||
      !first_number(ap.at("Exp"), &ape) || !equals_true(vc.at("Enable##Vorticity"), &vcOn) ||
      !first_number(vc.at("Coefficient##Vorticity"), &vcc) ||
      !first_number(vc.at("xSPH Viscosity Coefficient"), &xs))
    return false;
  // A negative count runs no iteration: a deliberate deviation, since the reference stores
  // it in a size_t (m_nbJacobiIters), where -1 wraps to SIZE_MAX and its iteration loop
  // would not end.
  const int32_t n = to_int(it);
  *iters = n > 0 ? (uint32_t)n : 0u;
  p->restDensity = (float)rho;
  p->relaxCFM = (float)cfm;
  p->timeStep = (float)dt;
  p->isArtPressureEnabled = apOn ? 1u : 0u;
  p->artPressureCoeff = (float)apc;
  p->artPressureRadius = (float)apr;
  p->artPressureExp = (uint32_t)to_int(ape);
  p->isVorticityConfEnabled = vcOn ? 1u : 0u;
  p->vorticityConfCoeff = (float)vcc;
  p->xsphViscosityCoeff = (float)xs;
  return true;
}

// "\"Fluids\": {...}" with the ranges a parameter widget shows (Fluids.cpp:52-74).
std::string fluids_tree(const rtp_fluid_params& p, uint32_t iters)
{
  return std::string("\"Fluids\": {") +
         "\"Rest Density\": " + range3(p.restDensity, 10.0f, 1000.0f) + ", " +
         "\"Relax CFM\": " + range3(p.relaxCFM, 100.0f, 1000.0f) + ", " +
         "\"Time Step\": " + range3(p.timeStep, 0.0001f, 0.020f) + ", " +
         "\"Nb Jacobi Iterations\": " + range3(iters, 1, 6) + ", " +
         "\"Artificial Pressure\": {\"Enable##Pressure\": " + (p.isArtPressureEnabled ? "true" : "false") +
         ", \"Coefficient##Pressure\": " + range3(p.artPressureCoeff, 0.0f, 0.001f) +
         ", \"Radius\": " + range3(p.artPressureRadius, 0.001f, 0.015f) +
         ", \"Exp\": " + range3(p.artPressureExp, 1, 6) + "}, " +
         "\"Vorticity Confinement\": {\"Enable##Vorticity\": " + (p.isVorticityConfEnabled ? "true" : "false") +
         ", \"Coefficient##Vorticity\": " + range3(p.vorticityConfCoeff, 0.0f, 0.001f) +
         ", \"xSPH Viscosity Coefficient\": " + range3(p.xsphViscosityCoeff, 0.0f, 0.001f) + "}}";
}

}  // namespace

extern "C" {

int rtp_fluid_params_from_json(const char* json, rtp_fluid_params* p, uint32_t* iters)
{
  if (!p || !iters) {
    g_jsonError = "NULL argument";
    return bad();
  }
  JVal root;
  if (!parse(json, root))
    return bad();
  return read_fluids(root, p, iters) ? RTP_OK : bad();
}

int rtp_fluid_params_to_json(const rtp_fluid_params* p, uint32_t iters, char* buf, size_t cap,
                             size_t* len)
{
  if (!p) {
    g_jsonError = "NULL argument";
    return bad();
  }
  return put(buf, cap, len, "{" + fluids_tree(*p, iters) + "}");
}

int rtp_cloud_params_from_json(const char* json, rtp_fluid_params* p, rtp_cloud_params* q,
                               uint32_t* iters)
{
  if (!p || !q || !iters) {
    g_jsonError = "NULL argument";
    return bad();
  }
  JVal root;
  if (!parse(json, root))
    return bad();
  rtp_fluid_params fp = *p;
  uint32_t it = 0;
  if (!read_fluids(root, &fp, &it))
    return bad();
  const Path c = Path{ &root, "" }.at("Clouds");
  bool smooth;
  double gh, bc, gc, alr, ptr, lhc, wind;
  if (!as_bool(c.at("Enable Temperature Smoothing"), &smooth) ||
      !first_number(c.at("Ground Heat Coefficient"), &gh) ||
      !first_number(c.at("Buoyancy Heat Coefficient"), &bc) ||
      !first_number(c.at("Gravity Coefficient"), &gc) ||
      !first_number(c.at("Adiabatic Lapse Rate"), &alr) ||
      !first_number(c.at("Phase Transition Rate"), &ptr) ||
      !first_number(c.at("Latent Heat Coefficient"), &lhc) ||
      !first_number(c.at("Wind Coefficient"), &wind))
    return bad();
  // Clouds.cpp:283-302: rest density, time step and relaxation come from the "Fluids" tree
  *p = fp;
  *iters = it;
  q->restDensity = fp.restDensity;
  q->timeStep = fp.timeStep;
  q->relaxCFM = fp.relaxCFM;
  q->isTempSmoothingEnabled = smooth ? 1u : 0u;
  q->groundHeatCoeff = (float)gh;
  q->buoyancyCoeff = (float)bc;
  q->gravCoeff = (float)gc;
  q->adiabaticLapseRate = (float)alr;
  q->phaseTransitionRate = (float)ptr;
  q->latentHeatCoeff = (float)lhc;
  q->windCoeff = (float)wind;
  return RTP_OK;
}

int rtp_cloud_params_to_json(const rtp_fluid_params* p, const rtp_cloud_params* q, uint32_t iters,
                             char* buf, size_t cap, size_t* len)
{
  if (!p || !q) {
    g_jsonError = "NULL argument";
    return bad();
  }
  // the ranges of initCloudsJson's "Clouds" tree (Clouds.cpp:88-97)
  const std::string s =
      "{" + fluids_tree(*p, iters) + ", \"Clouds\": {" +
      "\"Enable Temperature Smoothing\": " + (q->isTempSmoothingEnabled ? "true" : "false") + ", " +
      "\"Ground Heat Coefficient\": " + range3(q->groundHeatCoeff, 0.0f, 1000.0f) + ", " +
      "\"Buoyancy Heat Coefficient\": " + range3(q->buoyancyCoeff, 0.0f, 5.0f) + ", " +
      "\"Gravity Coefficient\": " + range3(q->gravCoeff, 0.0f, 0.1f) + ", " +
      "\"Adiabatic Lapse Rate\": " + range3(q->adiabaticLapseRate, 0.0f, 20.0f) + ", " +
      "\"Phase Transition Rate\": " + range3(q->phaseTransitionRate, 0.0f, 20.0f) + ", " +
      "\"Latent Heat Coefficient\": " + range3(q->latentHeatCoeff, 0.0f, 0.100f) + ", " +
      "\"Wind Coefficient\": " + range3(q->windCoeff, 0.0f, 1.0f) + "}}";
  return put(buf, cap, len, s);
}

int rtp_boids_params_from_json(const char* json, rtp_boids_params* p, rtp_target_params* t)
{
  if (!p || !t) {
    g_jsonError = "NULL argument";
    return bad();
  }
  JVal root;
  if (!parse(json, root))
    return bad();
  const Path b = Path{ &root, "" }.at("Boids");
  const Path al = b.at("Alignment"), co = b.at("Cohesion"), se = b.at("Separation"), tg = b.at("Target");
  double vel, als, cos, ses, rad, attract;
  bool alOn, coOn, seOn, tOn, show;
  if (!first_number(b.at("Velocity"), &vel) || !as_bool(al.at("Enable##Alignment"), &alOn) ||
      !as_bool(se.at("Enable##Separation"), &seOn) || !as_bool(co.at("Enable##Cohesion"), &coOn) ||
      !as_bool(tg.at("Enable##Target"), &tOn) || !as_bool(tg.at("Show"), &show) ||
      !first_number(tg.at("Radius"), &rad) || !as_number(tg.at("Attract"), &attract))
    return bad();
  // a disabled rule's scale is not read (Boids.cpp:189-191)
  als = cos = ses = 0.0;
  if ((alOn && !first_number(al.at("Scale##Alignment"), &als)) ||
      (seOn && !first_number(se.at("Scale##Separation"), &ses)) ||
      (coOn && !first_number(co.at("Scale##Cohesion"), &cos)))
    return bad();
  p->velocityScale = (float)vel;
  p->alignmentScale = alOn ? (float)als : 0.0f;
  p->separationScale = seOn ? (float)ses : 0.0f;
  p->cohesionScale = coOn ? (float)cos : 0.0f;
  // Boids.cpp:196-203 and Target::show (physics/utils/Target.hpp): "Attract" becomes the
  // sign as an integer (true 1, false 0), the target shows only while active
  t->active = tOn ? 1 : 0;
  t->visible = (tOn && show) ? 1 : 0;
  t->radius = (float)rad;
  t->sign = to_int(attract);
  return RTP_OK;
}

int rtp_boids_params_to_json(const rtp_boids_params* p, const rtp_target_params* t, char* buf,
                             size_t cap, size_t* len)
{
  if (!p || !t) {
    g_jsonError = "NULL argument";
    return bad();
  }
  // the ranges of initBoidsJson (Boids.cpp:41-73); every rule is written enabled with its
  // scale, so a zero scale reads back as zero; "Attract" is a boolean there, so a sign
  // other than 1 or 0 does not survive the round trip
  const std::string s =
      std::string("{\"Boids\": {") +
      "\"Velocity\": " + range3(p->velocityScale, 0.01f, 5.0f) + ", " +
      "\"Target\": {\"Enable##Target\": " + (t->active ? "true" : "false") +
      ", \"Show\": " + (t->visible ? "true" : "false") +
      ", \"Radius\": " + range3(t->radius, 1.0f, 20.0f) +
      ", \"Attract\": " + (t->sign > 0 ? "true" : "false") + "}, " +
      "\"Alignment\": {\"Enable##Alignment\": true, \"Scale##Alignment\": " + range3(p->alignmentScale, 0.0f, 3.0f) + "}, " +
      "\"Cohesion\": {\"Enable##Cohesion\": true, \"Scale##Cohesion\": " + range3(p->cohesionScale, 0.0f, 3.0f) + "}, " +
      "\"Separation\": {\"Enable##Separation\": true, \"Scale##Separation\": " + range3(p->separationScale, 0.0f, 3.0f) + "}}}";
  return put(buf, cap, len, s);
}

int rtp_json_equal(const char* a, const char* b, int* equal)
{
  if (!equal) {
    g_jsonError = "NULL argument";
    return bad();
  }
  JVal x, y;
  if (!parse(a, x) || !parse(b, y))
    return bad();
  *equal = same(x, y) ? 1 : 0;
  return RTP_OK;
}

int rtp_json_merge_patch(const char* target, const char* patch, char* buf, size_t cap, size_t* len)
{
  JVal t, q;
  if (!parse(target, t) || !parse(patch, q))
    return bad();
  merge_patch(t, q);
  std::string out;
  write(out, t);
  return put(buf, cap, len, out);
}

}  // extern "C"
