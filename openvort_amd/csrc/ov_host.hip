// ov_host.hip — host-side configuration and frame I/O (include/openvort_host.h).
//
// A libconfig-syntax reader (the subset the reference's configuration files
// use, with libconfig's typing: no auto-conversion between integers and
// floats), the settings of the reference's load_conf / setup_init
// (src/configuration.c:169-535) applied to an ov_config, the reference's flow
// table by name (src/external_velocity.c:85-174), the boundary presets
// (src/boundary_images.c:26-163), and frames in save_tangle's format
// (src/vortex_utils.c:267-387).  No device code.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/openvort_host.h"

namespace ovhost {

// ---- configuration values ----------------------------------------------------

enum VType { V_INT, V_INT64, V_FLOAT, V_BOOL, V_STRING, V_ARRAY, V_LIST, V_GROUP };

struct Value {
  VType type = V_GROUP;
  long long i = 0;
  double f = 0;
  std::string s;
  std::vector<Value> elems;                           // array / list elements
  std::vector<std::pair<std::string, Value>> members; // group settings, in file order
  int line = 0;
};

// Recursive-descent reader of libconfig's grammar:
//   file     := setting*
//   setting  := name ('=' | ':') value (';' | ',')?
//   value    := scalar | '[' scalars? ']' | '(' values? ')' | '{' setting* '}'
//   scalar   := bool | int | int64 | hex | float | string+
// with '#', '//' and '/* */' comments.
class Reader {
 public:
  explicit Reader(const char* text) : p_(text) {}
  bool parse_file(Value* root) {
    root->type = V_GROUP;
    root->line = 1;
    while (true) {
      skip();
      if (!*p_) return true;
      if (!parse_setting(root)) return false;
    }
  }
  const std::string& error() const { return err_; }

 private:
  const char* p_;
  int line_ = 1;
  std::string err_;

  bool fail(const char* what) {
    char b[256];
    snprintf(b, sizeof b, "line %d: %s", line_, what);
    err_ = b;
    return false;
  }
  void skip() {
    while (*p_) {
      if (*p_ == '\n') {
        ++line_;
        ++p_;
      } else if (isspace((unsigned char)*p_)) {
        ++p_;
      } else if (*p_ == '#' || (p_[0] == '/' && p_[1] == '/')) {
        while (*p_ && *p_ != '\n') ++p_;
      } else if (p_[0] == '/' && p_[1] == '*') {
        p_ += 2;
        while (*p_ && !(p_[0] == '*' && p_[1] == '/')) {
          if (*p_ == '\n') ++line_;
          ++p_;
        }
        if (*p_) p_ += 2;
      } else {
        break;
      }
    }
  }
  static bool name_start(char c) { return isalpha((unsigned char)c) || c == '*'; }
  static bool name_char(char c) { return isalnum((unsigned char)c) || c == '*' || c == '_' || c == '-'; }

  bool parse_setting(Value* group) {
    if (!name_start(*p_)) return fail("expected a setting name");
    const char* b = p_;
    while (name_char(*p_)) ++p_;
    std::string name(b, p_);
    skip();
    if (*p_ != '=' && *p_ != ':') return fail("expected '=' or ':' after a setting name");
    ++p_;
    skip();
    Value v;
    if (!parse_value(&v)) return false;
    for (auto& m : group->members)
      if (m.first == name) return fail(("duplicate setting name " + name).c_str());
    group->members.emplace_back(name, std::move(v));
    skip();
    if (*p_ == ';' || *p_ == ',') ++p_;
    return true;
  }

  bool parse_value(Value* v) {
    v->line = line_;
    if (*p_ == '[') return parse_seq(v, V_ARRAY, ']');
    if (*p_ == '(') return parse_seq(v, V_LIST, ')');
    if (*p_ == '{') {
      ++p_;
      v->type = V_GROUP;
      while (true) {
        skip();
        if (*p_ == '}') {
          ++p_;
          return true;
        }
        if (!*p_) return fail("unterminated group");
        if (!parse_setting(v)) return false;
      }
    }
    return parse_scalar(v);
  }

  bool parse_seq(Value* v, VType t, char close) {
    ++p_;
    v->type = t;
    skip();
    if (*p_ == close) {
      ++p_;
      return true;
    }
    while (true) {
      Value e;
      skip();
      if (t == V_ARRAY) {
        if (!parse_scalar(&e)) return false;
        if (!v->elems.empty() && !same_scalar_type(v->elems[0].type, e.type))
          return fail("mismatched element type in array");
      } else if (!parse_value(&e)) {
        return false;
      }
      v->elems.push_back(std::move(e));
      skip();
      if (*p_ == ',') {
        ++p_;
        continue;
      }
      if (*p_ == close) {
        ++p_;
        return true;
      }
      return fail(t == V_ARRAY ? "expected ',' or ']' in array" : "expected ',' or ')' in list");
    }
  }
  static bool same_scalar_type(VType a, VType b) {
    const bool ia = a == V_INT || a == V_INT64, ib = b == V_INT || b == V_INT64;
    return a == b || (ia && ib);
  }

  bool parse_string(std::string* out) {
    ++p_;
    while (*p_ && *p_ != '"') {
      if (*p_ == '\n') return fail("newline in a string");
      if (*p_ == '\\') {
        ++p_;
        switch (*p_) {
          case '\\': out->push_back('\\'); break;
          case '"': out->push_back('"'); break;
          case 'n': out->push_back('\n'); break;
          case 'r': out->push_back('\r'); break;
          case 't': out->push_back('\t'); break;
          case 'f': out->push_back('\f'); break;
          case 'x': {
            if (!isxdigit((unsigned char)p_[1]) || !isxdigit((unsigned char)p_[2]))
              return fail("bad \\x escape");
            char h[3] = {p_[1], p_[2], 0};
            out->push_back((char)strtol(h, nullptr, 16));
            p_ += 2;
            break;
          }
          default: return fail("unknown escape in a string");
        }
        ++p_;
      } else {
        out->push_back(*p_++);
      }
    }
    if (*p_ != '"') return fail("unterminated string");
    ++p_;
    return true;
  }

  bool parse_scalar(Value* v) {
    v->line = line_;
    if (*p_ == '"') {
      v->type = V_STRING;
      // adjacent string literals are concatenated
      while (*p_ == '"') {
        if (!parse_string(&v->s)) return false;
        skip();
      }
      return true;
    }
    if (name_start(*p_)) {
      const char* b = p_;
      while (isalpha((unsigned char)*p_)) ++p_;
      std::string w(b, p_);
      for (auto& c : w) c = (char)tolower((unsigned char)c);
      if (w == "true" || w == "false") {
        v->type = V_BOOL;
        v->i = w == "true";
        return true;
      }
      return fail("expected a value");
    }
    // numbers: hex integers, decimal integers (64-bit with L or when they do
    // not fit 32 bits), floats (a '.' or an exponent)
    const char* b = p_;
    const char* q = p_;
    if (*q == '+' || *q == '-') ++q;
    if (q[0] == '0' && (q[1] == 'x' || q[1] == 'X')) {
      q += 2;
      const char* h = q;
      while (isxdigit((unsigned char)*q)) ++q;
      if (q == h) return fail("bad hexadecimal number");
      unsigned long long u = strtoull(h, nullptr, 16);
      bool big = false;
      if (*q == 'L') {
        big = true;
        ++q;
      }
      v->type = (big || u > 0xffffffffULL) ? V_INT64 : V_INT;
      // a 32-bit hexadecimal integer is its bit pattern as an int
      long long x = v->type == V_INT ? (long long)(int)(unsigned)u : (long long)u;
      v->i = *b == '-' ? -x : x;
      p_ = q;
      return true;
    }
    bool digits = false, dot = false, expo = false;
    while (isdigit((unsigned char)*q)) {
      ++q;
      digits = true;
    }
    if (*q == '.') {
      dot = true;
      ++q;
      while (isdigit((unsigned char)*q)) {
        ++q;
        digits = true;
      }
    }
    if (!digits) return fail("expected a value");
    if (*q == 'e' || *q == 'E') {
      const char* e = q + 1;
      if (*e == '+' || *e == '-') ++e;
      if (isdigit((unsigned char)*e)) {
        expo = true;
        while (isdigit((unsigned char)*e)) ++e;
        q = e;
      }
    }
    std::string num(b, q);
    if (dot || expo) {
      v->type = V_FLOAT;
      v->f = strtod(num.c_str(), nullptr);
      p_ = q;
      return true;
    }
    errno = 0;
    long long x = strtoll(num.c_str(), nullptr, 10);
    if (errno == ERANGE) return fail("integer out of range");
    bool big = false;
    if (*q == 'L') {
      big = true;
      ++q;
    }
    v->type = (big || x > 2147483647LL || x < -2147483648LL) ? V_INT64 : V_INT;
    v->i = x;
    p_ = q;
    return true;
  }
};

// config_lookup: a path of setting names separated by '.', where "[n]"
// selects element n of a list or array.
const Value* lookup(const Value& root, const char* path) {
  const Value* cur = &root;
  std::string s(path);
  size_t pos = 0;
  while (pos <= s.size()) {
    size_t dot = s.find('.', pos);
    std::string part = s.substr(pos, dot == std::string::npos ? std::string::npos : dot - pos);
    if (part.empty()) return nullptr;
    if (part[0] == '[') {
      if (part.back() != ']' || (cur->type != V_LIST && cur->type != V_ARRAY)) return nullptr;
      const long k = strtol(part.c_str() + 1, nullptr, 10);
      if (k < 0 || k >= (long)cur->elems.size()) return nullptr;
      cur = &cur->elems[(size_t)k];
    } else {
      if (cur->type != V_GROUP) return nullptr;
      const Value* next = nullptr;
      for (auto& m : cur->members)
        if (m.first == part) next = &m.second;
      if (!next) return nullptr;
      cur = next;
    }
    if (dot == std::string::npos) break;
    pos = dot + 1;
  }
  return cur;
}

// config_lookup_int / _float / _bool / _string without auto-conversion: only a
// setting of that type counts, anything else leaves the target unchanged.
bool get_int(const Value& g, const char* path, int* out) {
  const Value* v = lookup(g, path);
  if (!v || v->type != V_INT) return false;
  *out = (int)v->i;
  return true;
}
bool get_float(const Value& g, const char* path, double* out) {
  const Value* v = lookup(g, path);
  if (!v || v->type != V_FLOAT) return false;
  *out = v->f;
  return true;
}
bool get_bool(const Value& g, const char* path, int* out) {
  const Value* v = lookup(g, path);
  if (!v || v->type != V_BOOL) return false;
  *out = (int)v->i;
  return true;
}
bool get_string(const Value& g, const char* path, std::string* out) {
  const Value* v = lookup(g, path);
  if (!v || v->type != V_STRING) return false;
  *out = v->s;
  return true;
}
// config_setting_get_float_elem: a float element, 0 for anything else.
double float_elem(const Value* v, int k) {
  if (!v || (v->type != V_ARRAY && v->type != V_LIST) || k >= (int)v->elems.size()) return 0.0;
  const Value& e = v->elems[(size_t)k];
  return e.type == V_FLOAT ? e.f : 0.0;
}
// load_conf_vector (configuration.c:63-77): a three-element array.
bool get_vector(const Value& g, const char* path, double out[3]) {
  const Value* v = lookup(g, path);
  if (!v || v->type != V_ARRAY) return false;
  for (int i = 0; i < 3; ++i) out[i] = float_elem(v, i);
  return true;
}

// ---- the reference's flow table (external_velocity.c:85-174) ---------------------

struct FlowParam {
  const char* name;
  int vector;   // 1: a 3-vector ("external_v" -> v, "wave" -> w, "polarization" -> q)
};
struct FlowEntry {
  const char* name;
  int type;
  int n_params;                // the table's n_params (rotosc-boundary lists 3, counts 2)
  FlowParam params[4];
};
// The entries keep the table's parameter names and order: parameters are
// matched by name, and scalar parameters fill ov_flow.a[] in this order.
const FlowEntry FLOWS[] = {
    {"no flow", OV_FLOW_NONE, 0, {}},
    {"spherical", OV_FLOW_SPHERICAL, 2, {{"strength", 0}, {"cutoff", 0}}},
    {"simple", OV_FLOW_SIMPLE, 1, {{"external_v", 1}}},
    {"spherical and simple", OV_FLOW_SPHERICAL_AND_SIMPLE, 3,
     {{"external_v", 1}, {"strength", 0}, {"cutoff", 0}}},
    {"cylindrical", OV_FLOW_CYLINDRICAL, 2, {{"strength", 0}, {"cutoff", 0}}},
    {"oscillating", OV_FLOW_OSCILLATING, 4,
     {{"strength", 0}, {"frequency", 0}, {"wave", 1}, {"polarization", 1}}},
    {"oscillating slit", OV_FLOW_OSCILLATING_SLIT, 3, {{"strength", 0}, {"frequency", 0}, {"cutoff", 0}}},
    {"coscos", OV_FLOW_COSCOS, 3, {{"k", 0}, {"v0", 0}, {"l", 0}}},
    {"coscos-divfree", OV_FLOW_COSCOS_DIVFREE, 3, {{"k", 0}, {"v0", 0}, {"l", 0}}},
    {"cos-divfree", OV_FLOW_COS_DIVFREE, 3, {{"k", 0}, {"v0", 0}, {"l", 0}}},
    {"simple-shear", OV_FLOW_SIMPLE_SHEAR, 2, {{"rate", 0}, {"v0", 0}}},
    {"rotosc", OV_FLOW_ROTOSC, 2, {{"freq", 0}, {"amp", 0}}},
    {"DG", OV_FLOW_DG, 2, {{"rot", 0}, {"vz", 0}}},
    {"rotosc-boundary", OV_FLOW_ROTOSC_BOUNDARY, 2, {{"freq", 0}, {"amp", 0}, {"delta", 0}}},
    {"constant acceleration", OV_FLOW_CONSTANT_ACCELERATION, 1, {{"accel", 0}}},
    {"parabolic", OV_FLOW_PARABOLIC, 2, {{"H", 0}, {"amp", 0}}},
};
constexpr int N_FLOWS = (int)(sizeof FLOWS / sizeof FLOWS[0]);

const FlowEntry* find_flow(const char* name) {
  for (int k = 0; k < N_FLOWS; ++k)
    if (strcmp(FLOWS[k].name, name) == 0) return &FLOWS[k];
  return nullptr;
}

double* vector_slot(ov_flow* f, const char* name) {
  if (strcmp(name, "external_v") == 0) return f->v;
  if (strcmp(name, "wave") == 0) return f->w;
  return f->q;   // "polarization"
}

void clear_flow(ov_flow* f, int type) {
  memset(f, 0, sizeof *f);
  f->type = type;
}

// setup_external_velocity (configuration.c:125-167) on one flow group.
bool flow_from_group(const Value& g, ov_flow* out, std::string* err) {
  std::string type;
  if (!get_string(g, "type", &type)) {
    *err = "line " + std::to_string(g.line) + ": flow group without a type string";
    return false;
  }
  const FlowEntry* e = find_flow(type.c_str());
  if (!e) {
    *err = "line " + std::to_string(g.line) + ": unknown flow type \"" + type + "\"";
    return false;
  }
  clear_flow(out, e->type);
  int scalar = 0;
  for (int k = 0; k < e->n_params; ++k) {
    const FlowParam& fp = e->params[k];
    if (fp.vector) {
      const Value* v = lookup(g, fp.name);
      double* dst = vector_slot(out, fp.name);
      for (int i = 0; i < 3; ++i) dst[i] = float_elem(v, i);
    } else {
      double x;
      if (!get_float(g, fp.name, &x)) {
        *err = "line " + std::to_string(g.line) + ": could not find parameter " + fp.name;
        return false;
      }
      out->a[scalar] = x;
    }
    if (!fp.vector) ++scalar;
  }
  return true;
}

// ---- boundary presets (boundary_images.c:26-163, configuration.c:322-386) ---------

struct Img {
  int s[3];
  int reflect;
};
struct Preset {
  const char* name;
  int n;                  // images used (the reference's .n, which for "wall-2-26" is 6)
  const Img* images;
  int walls[6];           // X_L X_H Y_L Y_H Z_L Z_H
};
constexpr int P = OV_WALL_PERIODIC, M = OV_WALL_MIRROR, O = OV_WALL_OPEN;
const Img IMG_PZ[] = {{{0, 0, -1}, -1}, {{0, 0, 1}, -1}};
const Img IMG_P6[] = {{{-1, 0, 0}, -1}, {{0, -1, 0}, -1}, {{0, 0, -1}, -1},
                      {{0, 0, 1}, -1},  {{0, 1, 0}, -1},  {{1, 0, 0}, -1}};
const Img IMG_P18[] = {{{-1, -1, 0}, -1}, {{-1, 0, -1}, -1}, {{-1, 0, 0}, -1}, {{-1, 0, 1}, -1},
                       {{-1, 1, 0}, -1},  {{0, -1, -1}, -1}, {{0, -1, 0}, -1}, {{0, -1, 1}, -1},
                       {{0, 0, -1}, -1},  {{0, 0, 1}, -1},   {{0, 1, -1}, -1}, {{0, 1, 0}, -1},
                       {{0, 1, 1}, -1},   {{1, -1, 0}, -1},  {{1, 0, -1}, -1}, {{1, 0, 0}, -1},
                       {{1, 0, 1}, -1},   {{1, 1, 0}, -1}};
const Img IMG_P26[] = {{{-1, -1, -1}, -1}, {{-1, -1, 0}, -1}, {{-1, -1, 1}, -1}, {{-1, 0, -1}, -1},
                       {{-1, 0, 0}, -1},   {{-1, 0, 1}, -1},  {{-1, 1, -1}, -1}, {{-1, 1, 0}, -1},
                       {{-1, 1, 1}, -1},   {{0, -1, -1}, -1}, {{0, -1, 0}, -1},  {{0, -1, 1}, -1},
                       {{0, 0, -1}, -1},   {{0, 0, 1}, -1},   {{0, 1, -1}, -1},  {{0, 1, 0}, -1},
                       {{0, 1, 1}, -1},    {{1, -1, -1}, -1}, {{1, -1, 0}, -1},  {{1, -1, 1}, -1},
                       {{1, 0, -1}, -1},   {{1, 0, 0}, -1},   {{1, 0, 1}, -1},   {{1, 1, -1}, -1},
                       {{1, 1, 0}, -1},    {{1, 1, 1}, -1}};
const Img IMG_W24[] = {{{-1, 0, 0}, -1}, {{1, 0, 0}, -1},       {{0, -1, 0}, -1},
                       {{0, 1, 0}, -1},  {{0, 0, -1}, OV_Z_L}, {{0, 0, 1}, OV_Z_H}};
// the first six entries of the reference's 26-entry list (its .n is 6)
const Img IMG_W226[] = {{{-1, 0, 0}, -1}, {{1, 0, 0}, -1},  {{0, -1, 0}, -1},
                        {{0, 1, 0}, -1},  {{-1, -1, 0}, -1}, {{-1, 1, 0}, -1}};
const Img IMG_W22[] = {{{-1, 0, 0}, -1},      {{1, 0, 0}, -1},       {{0, -1, 0}, OV_Y_L},
                       {{0, 1, 0}, OV_Y_H},   {{0, 0, -1}, OV_Z_L}, {{0, 0, 1}, OV_Z_H}};
const Img IMG_W1O[] = {{{0, 0, -1}, OV_Z_L}};
const Img IMG_W16[] = {{{-1, 0, 0}, -1}, {{0, -1, 0}, -1}, {{0, 0, -1}, OV_Z_L}, {{0, 1, 0}, -1},
                       {{1, 0, 0}, -1}};
const Img IMG_W118[] = {{{-1, -1, 0}, -1},     {{-1, 0, -1}, OV_Z_L}, {{-1, 0, 0}, -1},
                        {{-1, 0, 1}, -1},      {{-1, 1, 0}, -1},      {{0, -1, -1}, OV_Z_L},
                        {{0, -1, 0}, -1},      {{0, -1, 1}, -1},      {{0, 0, -1}, OV_Z_L},
                        {{0, 0, 1}, -1},       {{0, 1, -1}, OV_Z_L}, {{0, 1, 0}, -1},
                        {{0, 1, 1}, -1},       {{1, -1, 0}, -1},      {{1, 0, -1}, OV_Z_L},
                        {{1, 0, 0}, -1},       {{1, 0, 1}, -1},       {{1, 1, 0}, -1}};
const Img IMG_W126[] = {{{-1, -1, -1}, OV_Z_L}, {{-1, -1, 0}, -1}, {{-1, -1, 1}, -1},
                        {{-1, 0, -1}, OV_Z_L},  {{-1, 0, 0}, -1},  {{-1, 0, 1}, -1},
                        {{-1, 1, -1}, OV_Z_L},  {{-1, 1, 0}, -1},  {{-1, 1, 1}, -1},
                        {{0, -1, -1}, OV_Z_L},  {{0, -1, 0}, -1},  {{0, -1, 1}, -1},
                        {{0, 0, -1}, OV_Z_L},   {{0, 0, 1}, -1},   {{0, 1, -1}, OV_Z_L},
                        {{0, 1, 0}, -1},        {{0, 1, 1}, -1},   {{1, -1, -1}, OV_Z_L},
                        {{1, -1, 0}, -1},       {{1, -1, 1}, -1},  {{1, 0, -1}, OV_Z_L},
                        {{1, 0, 0}, -1},        {{1, 0, 1}, -1},   {{1, 1, -1}, OV_Z_L},
                        {{1, 1, 0}, -1},        {{1, 1, 1}, -1}};
// the reference lists (0, 0, 1) twice for "channel-z"
const Img IMG_CZ[] = {{{-1, 0, 0}, OV_X_L}, {{1, 0, 0}, OV_X_H}, {{0, -1, 0}, OV_Y_L},
                      {{0, 1, 0}, OV_Y_H},  {{0, 0, 1}, -1},     {{0, 0, 1}, -1}};

const Preset PRESETS[] = {
    {"wall-2-4", 6, IMG_W24, {P, P, P, P, M, M}},
    {"wall-2-2", 6, IMG_W22, {P, P, M, M, M, M}},
    {"wall-2-26", 6, IMG_W226, {P, P, P, P, M, M}},
    {"periodic-6", 6, IMG_P6, {P, P, P, P, P, P}},
    {"periodic-18", 18, IMG_P18, {P, P, P, P, P, P}},
    {"periodic-26", 26, IMG_P26, {P, P, P, P, P, P}},
    {"periodic-z-open-xy", 2, IMG_PZ, {O, O, O, O, P, P}},
    {"wall-1-6", 5, IMG_W16, {P, P, P, P, M, O}},
    {"wall-1-18", 18, IMG_W118, {P, P, P, P, M, O}},
    {"wall-1-26", 26, IMG_W126, {P, P, P, P, M, O}},
    {"wall-1-open", 1, IMG_W1O, {O, O, O, O, M, O}},
    {"open", 0, nullptr, {O, O, O, O, O, O}},
    {"channel-z", 6, IMG_CZ, {M, M, M, M, P, P}},
};

}  // namespace ovhost

using namespace ovhost;

namespace {

void put_err(char* err, int errlen, const std::string& msg) {
  if (err && errlen > 0) {
    strncpy(err, msg.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = 0;
  }
}

bool read_file(const char* path, std::string* out) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  char buf[65536];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) out->append(buf, k);
  fclose(f);
  return true;
}

// save_point's columns (vortex_utils.c:272-285)
void write_point(FILE* s, int vort, const ov_tangle_state* t, int i, const char* g) {
  auto vec = [&](const ov_vec3d* a) {
    fprintf(s, g, a[i].p[0]);
    fputc('\t', s);
    fprintf(s, g, a[i].p[1]);
    fputc('\t', s);
    fprintf(s, g, a[i].p[2]);
  };
  fprintf(s, "%d\t", vort);
  vec(t->vnodes);
  fputc('\t', s);
  vec(t->vels);
  fputc('\t', s);
  vec(t->tangents);
  fputc('\t', s);
  vec(t->normals);
  fprintf(s, "\t%d\t%d\t%d\t%d\t%d\t", i, t->connections[i].reverse, t->connections[i].forward,
          t->status[i].status, t->status[i].pin_wall);
  fprintf(s, g, t->dxi[i]);
  fputc('\n', s);
}

}  // namespace

extern "C" {

void ov_config_default(ov_config* c) {
  memset(c, 0, sizeof *c);
  c->max_steps = -1;
  c->frame_shots = 100;
  c->kappa = 9.97e-4;
  c->num_threads = 4;
  c->pin_mode = OV_PINNED;
  c->use_BH = 0;
  c->BH_resolution = 0.1;
  c->BH_quadtree = 0;
  c->BH_grain = 1e-8;
  c->use_mutual_friction = 1;
  c->alpha = 0.1;
  c->alpha_p = 0.01;
  c->LIA_only = 0;
  c->hyperfriction = 1;
  c->hyperalpha = 0.5;
  c->max_curvature_scale = 0.2;
  c->dt = 1e-3;
  c->dl_min = 1e-3;
  c->dl_max = 5e-3;
  c->small_loop_cutoff = 5;
  c->reconnection_angle_cutoff = 0.087;
  c->rec_dist = 1e-3;
  c->eliminate_loops_origin_cutoff = 3e-2;
  c->eliminate_loops_zaxis_cutoff = 2e-2;
  c->eliminate_outer_loops_cutoff = 1;
  c->loop_injection_frequency = 1;
  c->line_injection_n = 1;
  c->vn.type = c->vs.type = c->vb.type = OV_FLOW_NONE;
  for (int i = 0; i < 3; ++i) {
    c->box.bottom_left_back.p[i] = 0;
    c->box.top_right_front.p[i] = 1;
  }
  for (int f = 0; f < 6; ++f) c->box.wall[f] = OV_WALL_OPEN;
}

int ov_config_parse(const char* text, ov_config* c, char* err, int errlen) {
  if (!text || !c) return OV_E_INVALID;
  Value root;
  Reader rd(text);
  if (!rd.parse_file(&root)) {
    put_err(err, errlen, rd.error());
    return OV_E_INVALID;
  }
  auto bad = [&](const std::string& m) {
    put_err(err, errlen, m);
    return OV_E_INVALID;
  };
  std::string str;
  int iv;
  double dv;
  // load_conf (configuration.c:183-300): general properties
  get_int(root, "max_steps", &c->max_steps);
  get_int(root, "frame_shots", &c->frame_shots);
  get_float(root, "KAPPA", &c->kappa);
  get_int(root, "num_threads", &c->num_threads);
  if (get_string(root, "wall_type", &str)) {
    if (str == "pin")
      c->pin_mode = OV_PINNED;
    else if (str == "slip")
      c->pin_mode = OV_PINNED_SLIP;
    else
      return bad("Unknown wall type. Possible options are 'pin' or 'slip'.");
  }
  get_bool(root, "use_BH", &c->use_BH);
  get_float(root, "BH_resolution", &c->BH_resolution);
  get_bool(root, "BH_quadtree", &c->BH_quadtree);
  get_float(root, "BH_grain", &c->BH_grain);
  get_bool(root, "use_mutual_friction", &c->use_mutual_friction);
  get_float(root, "alpha", &c->alpha);
  get_float(root, "alpha_p", &c->alpha_p);
  get_bool(root, "LIA", &c->LIA_only);
  get_bool(root, "use_hyperfriction", &c->hyperfriction);
  get_float(root, "hyperalpha", &c->hyperalpha);
  get_float(root, "max_curvature_scale", &c->max_curvature_scale);
  get_float(root, "dt", &c->dt);
  get_float(root, "dl_min", &c->dl_min);
  get_float(root, "dl_max", &c->dl_max);
  get_int(root, "small_loop_cutoff", &c->small_loop_cutoff);
  if (get_float(root, "reconnection_angle_cutoff", &dv)) c->reconnection_angle_cutoff = M_PI / 180.0 * dv;
  get_float(root, "reconnection_distance", &c->rec_dist);
  get_bool(root, "eliminate_origin_loops", &c->eliminate_origin_loops);
  get_float(root, "eliminate_loops_origin_cutoff", &c->eliminate_loops_origin_cutoff);
  get_bool(root, "eliminate_zaxis_loops", &c->eliminate_zaxis_loops);
  get_float(root, "eliminate_loops_zaxis_cutoff", &c->eliminate_loops_zaxis_cutoff);
  get_bool(root, "eliminate_outer_loops", &c->eliminate_outer_loops);
  get_float(root, "eliminate_outer_loops_cutoff", &c->eliminate_outer_loops_cutoff);
  if (get_bool(root, "loop_injection", &iv)) {
    c->loop_injection = iv;
    if (iv && !get_float(root, "loop_injection_frequency", &c->loop_injection_frequency))
      return bad("Specify loop_injection_frequency in the config.");
  }
  if (get_bool(root, "line_injection", &iv)) {
    c->line_injection = iv;
    if (iv) {
      if (!get_float(root, "line_injection_frequency", &c->line_injection_frequency))
        return bad("Specify line_injection_frequency in the config.");
      if (!get_int(root, "line_injection_n", &c->line_injection_n)) c->line_injection_n = 1;
      if (!get_bool(root, "line_injection_polarized", &c->line_injection_polarized))
        c->line_injection_polarized = 0;
    }
  }
  // external velocities (configuration.c:302-316); an absent vb_conf is "no flow"
  std::string e;
  struct {
    const char* key;
    ov_flow* f;
    int* set;
  } groups[3] = {{"vn_conf", &c->vn, &c->vn_set}, {"vs_conf", &c->vs, &c->vs_set},
                 {"vb_conf", &c->vb, &c->vb_set}};
  for (auto& g : groups) {
    const Value* v = lookup(root, g.key);
    if (v && v->type == V_GROUP) {
      if (!flow_from_group(*v, g.f, &e)) return bad(std::string(g.key) + ": " + e);
      *g.set = 1;
    } else if (g.f == &c->vb) {
      clear_flow(&c->vb, OV_FLOW_NONE);
    }
  }
  // domain box (configuration.c:318-328)
  const Value* dom = lookup(root, "domain");
  if (!dom || dom->type != V_LIST) return bad("Error: incorrectly specified domain");
  double lo[3], hi[3];
  if (!get_vector(root, "domain.[0]", lo) || !get_vector(root, "domain.[1]", hi))
    return bad("Failed to load vector: domain");
  for (int i = 0; i < 3; ++i) {
    c->box.bottom_left_back.p[i] = lo[i];
    c->box.top_right_front.p[i] = hi[i];
  }
  // boundaries (configuration.c:330-387)
  if (!get_string(root, "boundaries", &str)) return bad("Error: specify boundaries");
  const Preset* pr = nullptr;
  for (const Preset& q : PRESETS)
    if (str == q.name) pr = &q;
  if (!pr) return bad("Error: unknown boundary condition \"" + str + "\"");
  snprintf(c->boundaries, sizeof c->boundaries, "%s", pr->name);
  c->n_images = pr->n;
  for (int j = 0; j < pr->n; ++j) {
    for (int i = 0; i < 3; ++i) c->images[j].shift[i] = pr->images[j].s[i];
    c->images[j].reflect = pr->images[j].reflect;
  }
  for (int f = 0; f < 6; ++f) c->box.wall[f] = pr->walls[f];
  // the initial condition (setup_init, configuration.c:407-535): its mode and
  // the parameters that mode needs
  if (!get_string(root, "init_mode", &str)) return bad("Error: set the initialization mode");
  snprintf(c->init_mode, sizeof c->init_mode, "%s", str.c_str());
  auto need_int = [&](const char* k, int* dst) { return get_int(root, k, dst); };
  auto need_float = [&](const char* k, double* dst) { return get_float(root, k, dst); };
  bool ok = true;
  if (str == "random loops") {
    ok = need_int("init_loops_n", &c->init_loops_n);
  } else if (str == "one loop") {
    get_vector(root, "loop_center", c->loop_center);
    get_vector(root, "loop_dir", c->loop_dir);
    ok = need_float("loop_r", &c->loop_r) && need_int("loop_N", &c->loop_N);
  } else if (str == "restart") {
    std::string path;
    ok = get_string(root, "init_file", &path);
    snprintf(c->init_file, sizeof c->init_file, "%s", path.c_str());
  } else if (str == "big ring") {
    ok = need_float("ring_r", &c->ring_r) && need_int("ring_N", &c->ring_N);
  } else if (str == "random straight lines") {
    ok = need_int("init_npairs", &c->init_npairs) && need_int("init_line_points", &c->init_line_points);
  } else if (str == "random polarized lines") {
    ok = need_int("init_nvorts", &c->init_nvorts) && need_int("init_line_points", &c->init_line_points) &&
         need_int("init_direction", &c->init_direction);
  } else if (str == "lattice") {
    ok = need_int("init_lattice_shells", &c->init_lattice_shells) &&
         need_int("init_line_points", &c->init_line_points) &&
         need_int("init_direction", &c->init_direction) &&
         need_float("init_lattice_constant", &c->init_lattice_constant);
  } else if (str == "quad") {
    ok = need_int("init_line_points", &c->init_line_points) &&
         need_float("init_separation", &c->init_separation);
  } else if (str == "dipole") {
    ok = need_int("init_line_points", &c->init_line_points) &&
         need_int("dipole_direction", &c->dipole_direction) &&
         need_float("init_separation", &c->init_separation);
  } else if (str == "single") {
    ok = need_int("init_line_points", &c->init_line_points) &&
         need_int("init_line_direction", &c->init_line_direction);
    if (ok && !get_int(root, "init_k_KW", &c->init_k_KW)) c->init_k_KW = 0;
    if (ok && !get_float(root, "init_r_KW", &c->init_r_KW)) c->init_r_KW = 0;
  } else {
    return bad("Error: unknown initialization mode: " + str);
  }
  if (!ok) return bad("Failed to setup initial condition (" + str + ")");
  return OV_OK;
}

int ov_config_load(const char* path, ov_config* c, char* err, int errlen) {
  std::string text;
  if (!path || !read_file(path, &text)) {
    put_err(err, errlen, std::string("cannot read ") + (path ? path : "(null)"));
    return OV_E_INVALID;
  }
  return ov_config_parse(text.c_str(), c, err, errlen);
}

void ov_params_from_config(const ov_config* c, ov_params* p) {
  p->kappa = c->kappa;
  p->alpha = c->alpha;
  p->alpha_p = c->alpha_p;
  p->dl_max = c->dl_max;
  p->LIA_only = c->LIA_only;
  p->use_mutual_friction = c->use_mutual_friction;
  p->hyperfriction = c->hyperfriction;
  p->use_BH = c->use_BH;
  p->max_curvature_scale = c->max_curvature_scale;
  p->hyperalpha = c->hyperalpha;
  p->BH_resolution = c->BH_resolution;
  p->BH_grain = c->BH_grain;
  p->BH_quadtree = c->BH_quadtree;
  p->vn = c->vn;
  p->vs = c->vs;
  p->vb = c->vb;
}

int ov_flow_from_vconf(const ov_v_conf* conf, ov_flow* out) {
  if (!conf || !conf->name || !out) return OV_E_INVALID;
  const FlowEntry* e = find_flow(conf->name);
  if (!e) return OV_E_INVALID;
  clear_flow(out, e->type);
  // get_v_param_scalar / get_v_param_vector: by name over the first n_params
  const int n = conf->n_params < 32 ? conf->n_params : 32;
  int scalar = 0;
  for (int k = 0; k < e->n_params; ++k) {
    const FlowParam& fp = e->params[k];
    for (int j = 0; j < n; ++j) {
      const ov_v_param& q = conf->v_params[j];
      if (!q.name || strcmp(q.name, fp.name) != 0) continue;
      if (fp.vector) {
        double* dst = vector_slot(out, fp.name);
        for (int i = 0; i < 3; ++i) dst[i] = q.value.vector.p[i];
      } else {
        out->a[scalar] = q.value.scalar;
      }
      break;
    }
    if (!fp.vector) ++scalar;
  }
  return OV_OK;
}

int ov_frame_save(const char* path, const ov_tangle_state* t, int digits) {
  if (!path || !t || t->N < 0) return OV_E_INVALID;
  if (t->N > 0 && (!t->vnodes || !t->vels || !t->tangents || !t->normals || !t->dxi ||
                   !t->connections || !t->status))
    return OV_E_INVALID;
  FILE* s = fopen(path, "w");
  if (!s) return OV_E_INVALID;
  const char* g = digits >= 17 ? "%.17g" : "%.15g";
  // the walk of save_tangle (vortex_utils.c:306-352): every unvisited
  // non-EMPTY point starts a vortex, followed forward; an open end (forward
  // -1) continues backwards from the start's reverse neighbour
  std::vector<char> visited((size_t)t->N, 0);
  int vortex = 0;
  for (int k = 0; k < t->N; ++k) {
    if (visited[(size_t)k]) continue;
    if (t->status[k].status == OV_EMPTY) {
      visited[(size_t)k] = 1;
      continue;
    }
    const int first = k;
    int curr = k;
    while (t->connections[curr].forward != first) {
      write_point(s, vortex, t, curr, g);
      visited[(size_t)curr] = 1;
      curr = t->connections[curr].forward;
      if (curr < 0) {
        curr = t->connections[first].reverse;
        while (curr > 0 && t->connections[curr].reverse > 0) {
          write_point(s, vortex, t, curr, g);
          visited[(size_t)curr] = 1;
          curr = t->connections[curr].reverse;
        }
        break;
      }
    }
    if (curr > 0) {
      write_point(s, vortex, t, curr, g);
      visited[(size_t)curr] = 1;
    }
    ++vortex;
  }
  const bool ok = !ferror(s);
  fclose(s);
  return ok ? OV_OK : OV_E_INVALID;
}

namespace {
struct FrameRow {
  int vort;
  double v[12];
  int ncol;
  int idx, rev, fwd, status, pin;
  double dxi;
};

int read_frame(const char* path, std::vector<FrameRow>* rows, std::string* err) {
  std::string text;
  if (!read_file(path, &text)) {
    *err = std::string("cannot read ") + path;
    return OV_E_INVALID;
  }
  size_t pos = 0;
  int line = 0;
  while (pos < text.size()) {
    size_t end = text.find('\n', pos);
    if (end == std::string::npos) end = text.size();
    std::string ln = text.substr(pos, end - pos);
    pos = end + 1;
    ++line;
    std::vector<std::string> tok;
    const char* q = ln.c_str();
    while (*q) {
      while (*q && isspace((unsigned char)*q)) ++q;
      if (!*q) break;
      const char* b = q;
      while (*q && !isspace((unsigned char)*q)) ++q;
      tok.emplace_back(b, q);
    }
    if (tok.empty()) continue;
    if (tok.size() != 13 && tok.size() != 19) {
      *err = "line " + std::to_string(line) + ": " + std::to_string(tok.size()) +
             " columns (13 or 19 expected)";
      return OV_E_INVALID;
    }
    FrameRow r{};
    char* e = nullptr;
    r.vort = (int)strtol(tok[0].c_str(), &e, 10);
    bool ok = *e == 0;
    for (int c = 0; c < 12; ++c) {
      r.v[c] = strtod(tok[(size_t)c + 1].c_str(), &e);
      ok = ok && *e == 0;
    }
    r.ncol = (int)tok.size();
    if (r.ncol == 19) {
      int* ints[5] = {&r.idx, &r.rev, &r.fwd, &r.status, &r.pin};
      for (int c = 0; c < 5; ++c) {
        *ints[c] = (int)strtol(tok[(size_t)c + 13].c_str(), &e, 10);
        ok = ok && *e == 0;
      }
      r.dxi = strtod(tok[18].c_str(), &e);
      ok = ok && *e == 0;
    }
    if (!ok) {
      *err = "line " + std::to_string(line) + ": not a number";
      return OV_E_INVALID;
    }
    if (!rows->empty() && rows->back().ncol != r.ncol) {
      *err = "line " + std::to_string(line) + ": mixed 13- and 19-column rows";
      return OV_E_INVALID;
    }
    rows->push_back(r);
  }
  return OV_OK;
}
}  // namespace

int ov_frame_count(const char* path, int* n_points) {
  if (!path || !n_points) return OV_E_INVALID;
  std::vector<FrameRow> rows;
  std::string e;
  const int rc = read_frame(path, &rows, &e);
  if (rc) return rc;
  *n_points = (int)rows.size();
  return OV_OK;
}

int ov_frame_load(const char* path, ov_tangle_state* t, char* err, int errlen) {
  if (!path || !t) return OV_E_INVALID;
  std::vector<FrameRow> rows;
  std::string e;
  int rc = read_frame(path, &rows, &e);
  if (rc) {
    put_err(err, errlen, e);
    return rc;
  }
  const int n = (int)rows.size();
  if (t->N < n) {
    put_err(err, errlen, "tangle capacity " + std::to_string(t->N) + " below the frame's " +
                             std::to_string(n) + " points");
    return OV_E_INVALID;
  }
  if (t->N > 0 && (!t->vnodes || !t->vels || !t->tangents || !t->normals || !t->dxi ||
                   !t->connections || !t->status)) {
    put_err(err, errlen, "null tangle arrays");
    return OV_E_INVALID;
  }
  // vortex indices start at 0 and increase by one from vortex to vortex
  for (int k = 0; k < n; ++k) {
    const int prev = k ? rows[(size_t)k - 1].vort : 0;
    if (rows[(size_t)k].vort != prev && rows[(size_t)k].vort != prev + 1) {
      put_err(err, errlen, "point " + std::to_string(k) + ": vortex index out of sequence");
      return OV_E_INVALID;
    }
  }
  for (int k = 0; k < n; ++k) {
    const FrameRow& r = rows[(size_t)k];
    for (int i = 0; i < 3; ++i) {
      t->vnodes[k].p[i] = r.v[i];
      t->vels[k].p[i] = r.v[3 + i];
      t->tangents[k].p[i] = r.v[6 + i];
      t->normals[k].p[i] = r.v[9 + i];
    }
    t->status[k].status = OV_FREE;
    t->status[k].pin_wall = -1;
    t->dxi[k] = 0;
  }
  bool linked = false;
  if (n > 0 && rows[0].ncol == 19) {
    // links as saved, through the saved point indices (each point once)
    std::vector<int> where;
    int max_idx = 0;
    for (const FrameRow& r : rows) max_idx = std::max(max_idx, std::max(r.idx, std::max(r.fwd, r.rev)));
    where.assign((size_t)max_idx + 1, -1);
    bool unique = true;
    for (int k = 0; k < n; ++k) {
      const int i = rows[(size_t)k].idx;
      if (i < 0 || where[(size_t)i] != -1) unique = false;
      else where[(size_t)i] = k;
    }
    auto map = [&](int i) { return i < 0 ? -1 : where[(size_t)i]; };
    bool complete = unique;
    for (int k = 0; k < n && complete; ++k) {
      const FrameRow& r = rows[(size_t)k];
      if ((r.fwd >= 0 && map(r.fwd) < 0) || (r.rev >= 0 && map(r.rev) < 0)) complete = false;
    }
    if (complete) {
      for (int k = 0; k < n; ++k) {
        const FrameRow& r = rows[(size_t)k];
        t->connections[k].forward = map(r.fwd);
        t->connections[k].reverse = map(r.rev);
        t->status[k].status = r.status;
        t->status[k].pin_wall = r.pin;
        t->dxi[k] = r.dxi;
      }
      linked = true;
    } else {
      for (int k = 0; k < n; ++k) {
        t->status[k].status = rows[(size_t)k].status;
        t->status[k].pin_wall = rows[(size_t)k].pin;
        t->dxi[k] = rows[(size_t)k].dxi;
      }
    }
  }
  if (!linked) {
    // load_tangle's linking (vortex_utils.c:354-387): consecutive points of
    // a vortex in file order, each vortex closed into a loop
    int start = 0;
    for (int k = 0; k <= n; ++k) {
      if (k == n || (k > start && rows[(size_t)k].vort != rows[(size_t)start].vort)) {
        for (int j = start; j < k; ++j) {
          t->connections[j].forward = j + 1 < k ? j + 1 : start;
          t->connections[j].reverse = j > start ? j - 1 : k - 1;
        }
        start = k;
      }
    }
  }
  for (int k = n; k < t->N; ++k) {
    t->status[k].status = OV_EMPTY;
    t->status[k].pin_wall = -1;
    t->connections[k].forward = -1;
    t->connections[k].reverse = -1;
  }
  t->next_free = n;
  t->total_free = t->N - n;
  return OV_OK;
}

}  // extern "C"
