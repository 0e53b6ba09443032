// refcpu.cpp — CPU restatement of krets' MNA solve path.
//
// TEST INFRASTRUCTURE ONLY: this is the checker for libkrets_gpu.so. Only the
// test suite, the smoke check and bench.py's cpu_baseline leg load it.
//
// What it restates (paths relative to the krets repository):
//   netlist subset   crates/krets-parser/src/parser.rs:27-154, elements/*.rs
//                    parse_* functions, utils.rs:32-78, models.rs:40-95,
//                    elements/voltage_source.rs:100-235
//   index numbering  crates/krets-parser/src/parser.rs:105-125
//   device maths     elements/diode.rs:24-71, voltage_source.rs:50-93,243-252,
//                    constants.rs:1-4, models/diode.rs:14-37
//   stamps           crates/krets-solver/src/stampable.rs:202-1050
//   loops            solver.rs:55-104, solver/op.rs, solver/dc.rs, solver/ac.rs,
//                    solver/transient.rs, analyses.rs:78-158
// The sparse LU stands in for faer's sp_lu (third-party faer 0.23.2, not
// vendored): a fill-reducing column order (minimum degree on A+A^T) and a
// left-looking Gilbert-Peierls factorisation with row partial pivoting,
// recomputed from scratch on every solve in COMPAT cost mode.
//
// Additions that krets does not have (clean mode) follow published methods:
// trapezoidal companion models, SPICE3 pnjlim junction limiting (Nagel,
// SPICE2 thesis 1975), divided-difference LTE step control for backward Euler,
// and the discrete adjoint of the fixed-step recurrence.
//
// Beyond krets' element set: the linear controlled sources of krets' README
// to-do list (README.md:56-62), restated from the SPICE element definitions
// (SPICE3 user's guide, section on linear dependent sources): E (VCVS),
// G (VCCS), F (CCCS), H (CCVS). Pinned by closed-form circuits in the tests.

#include "refcpu.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

using cd = std::complex<double>;

// constants.rs:1-4
const double K_BOLTZ = 1.380649e-23;
const double Q_ELEM = 1.602176634e-19;
const double TEMPERATURE = 300.0;
const double VT = K_BOLTZ * TEMPERATURE / Q_ELEM;
const double PI_ = 3.14159265358979323846;

struct RefError {
  int code;
  std::string msg;
};

thread_local std::string g_last_error;

enum ET { E_R, E_C, E_L, E_V, E_I, E_D, E_E, E_G, E_F, E_H };
const char ET_LETTER[] = {'R', 'C', 'L', 'V', 'I', 'D', 'E', 'G', 'F', 'H'};

struct Wave {
  int kind = 0;  // 0 none, 1 pulse, 2 sin
  double p[7] = {0, 0, 0, 0, 0, 0, 0};
};

struct Elem {
  ET t;
  std::string name, plus, minus, model;
  double value = 0.0;  // R ohm, C farad, L henry, V dc, I amp, D Is
  bool g2 = false;
  double ac = 0.0;
  Wave w;
  double n_emis = 1.0;
  // diode junction charge (opt-in parasitics): cjo, vj, m, fc, tt
  double jq[5] = {0.0, 1.0, 0.5, 0.5, 0.0};
  bool has_jq() const { return t == E_D && (jq[0] != 0.0 || jq[4] != 0.0); }
  std::string cplus, cminus, ctrl;  // E/G controlling nodes, F/H controlling element id
  int ip = -1, im = -1, ic = -1;    // lean indices
  int icp = -1, icm = -1, icb = -1;  // controlling node / branch indices
  std::string ident() const { return std::string(1, ET_LETTER[t]) + name; }
  bool is_g2() const {  // elements.rs:124-140 (E and H own a branch current, like V)
    switch (t) {
      case E_V: case E_L: case E_I: case E_E: case E_H: return true;
      case E_R: case E_C: return g2;
      default: return false;
    }
  }
};

struct Circuit {
  std::vector<Elem> el;
  std::unordered_map<std::string, int> index;
  std::vector<std::string> names;
  bool nonlinear = false;
};

// ------------------------------------------------------------------ parsing
struct Cur {
  const std::string* sp;
  size_t p;
  Cur(const std::string& str, size_t pos) : sp(&str), p(pos) {}
  const std::string& str() const { return *sp; }
};

bool is_sp(char c) { return c == ' ' || c == '\t'; }
bool is_an(char c) { return std::isalnum((unsigned char)c) || c == '_'; }

bool space1(Cur& c) {
  size_t q = c.p;
  while (q < c.str().size() && is_sp(c.str()[q])) q++;
  if (q == c.p) return false;
  c.p = q;
  return true;
}
void space0(Cur& c) {
  while (c.p < c.str().size() && is_sp(c.str()[c.p])) c.p++;
}
bool tag_nc(Cur& c, const char* t) {
  size_t n = std::strlen(t);
  if (c.p + n > c.str().size()) return false;
  for (size_t i = 0; i < n; i++)
    if (std::tolower((unsigned char)c.str()[c.p + i]) != std::tolower((unsigned char)t[i])) return false;
  c.p += n;
  return true;
}
bool tag_exact(Cur& c, const char* t) {
  size_t n = std::strlen(t);
  if (c.str().compare(c.p, n, t) != 0) return false;
  c.p += n;
  return true;
}
bool ident1(Cur& c, std::string& out) {  // utils.rs:67-69
  size_t q = c.p;
  while (q < c.str().size() && is_an(c.str()[q])) q++;
  if (q == c.p) return false;
  out = c.str().substr(c.p, q - c.p);
  c.p = q;
  return true;
}

// Rust's f64::from_str grammar (decimal, optional exponent, inf/nan).
bool rust_f64(const std::string& t, double& v) {
  std::string u;
  for (char ch : t) u += (char)std::tolower((unsigned char)ch);
  size_t i = 0;
  if (i < u.size() && (u[i] == '+' || u[i] == '-')) i++;
  std::string body = u.substr(i);
  if (body == "inf" || body == "infinity" || body == "nan") {
    v = std::strtod(u.c_str(), nullptr);
    return true;
  }
  size_t d0 = i;
  while (i < u.size() && std::isdigit((unsigned char)u[i])) i++;
  size_t nd = i - d0;
  size_t nf = 0;
  if (i < u.size() && u[i] == '.') {
    i++;
    size_t f0 = i;
    while (i < u.size() && std::isdigit((unsigned char)u[i])) i++;
    nf = i - f0;
  }
  if (nd + nf == 0) return false;
  if (i < u.size() && u[i] == 'e') {
    i++;
    if (i < u.size() && (u[i] == '+' || u[i] == '-')) i++;
    size_t e0 = i;
    while (i < u.size() && std::isdigit((unsigned char)u[i])) i++;
    if (i == e0) return false;
  }
  if (i != u.size()) return false;
  v = std::strtod(u.c_str(), nullptr);
  return true;
}

// utils.rs:32-64 parse_value
bool parse_value(const std::string& tok, double& v) {
  std::string u;
  for (char ch : tok) u += (char)std::toupper((unsigned char)ch);
  double mult = 1.0;
  std::string num = u;
  if (u.size() >= 3 && u.compare(u.size() - 3, 3, "MEG") == 0) {
    num = u.substr(0, u.size() - 3);
    mult = 1e6;
  } else if (!u.empty()) {
    char l = u.back();
    bool suf = true;
    switch (l) {
      case 'F': mult = 1e-15; break;
      case 'P': mult = 1e-12; break;
      case 'N': mult = 1e-9; break;
      case 'U': mult = 1e-6; break;
      case 'M': mult = 1e-3; break;
      case 'K': mult = 1e3; break;
      case 'G': mult = 1e9; break;
      case 'T': mult = 1e12; break;
      default: suf = false;
    }
    if (suf) num = u.substr(0, u.size() - 1);
  }
  double base;
  if (!rust_f64(num, base)) return false;
  v = base * mult;
  return true;
}

// utils.rs:72-78 value_parser: a token of chars outside " \t\r\n()".
bool value_tok(Cur& c, double& v) {
  size_t q = c.p;
  while (q < c.str().size() && !std::strchr(" \t\r\n()", c.str()[q])) q++;
  if (q == c.p) return false;
  if (!parse_value(c.str().substr(c.p, q - c.p), v)) return false;
  c.p = q;
  return true;
}

// nom::number::complete::double — recognize_float or nan/infinity/inf.
// Returns 0 = no match, 1 = match, -1 = hard failure (exponent without digits).
int nom_double(Cur& c, double& v) {
  const std::string& s = c.str();
  size_t i = c.p;
  if (i < s.size() && (s[i] == '+' || s[i] == '-')) i++;
  size_t d0 = i;
  bool ok = false;
  while (i < s.size() && std::isdigit((unsigned char)s[i])) i++;
  if (i > d0) {
    ok = true;
    if (i < s.size() && s[i] == '.') {
      i++;
      while (i < s.size() && std::isdigit((unsigned char)s[i])) i++;
    }
  } else if (i < s.size() && s[i] == '.') {
    size_t f0 = i + 1, j = f0;
    while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
    if (j > f0) {
      ok = true;
      i = j;
    }
  }
  if (ok) {
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
      size_t j = i + 1;
      if (j < s.size() && (s[j] == '+' || s[j] == '-')) j++;
      size_t e0 = j;
      while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
      if (j == e0) return -1;
      i = j;
    }
    v = std::strtod(s.substr(c.p, i - c.p).c_str(), nullptr);
    c.p = i;
    return 1;
  }
  Cur t{s, c.p};
  if (tag_nc(t, "nan")) { v = NAN; c.p = t.p; return 1; }
  t.p = c.p;
  if (tag_nc(t, "infinity")) { v = INFINITY; c.p = t.p; return 1; }
  t.p = c.p;
  if (tag_nc(t, "inf")) { v = INFINITY; c.p = t.p; return 1; }
  return 0;
}

bool two_nodes(Cur& c, Elem& e) {
  return ident1(c, e.name) && space1(c) && ident1(c, e.plus) && space1(c) && ident1(c, e.minus);
}

bool parse_rlci(const std::string& line, char letter, ET t, Elem& e) {
  Cur c{line, 0};
  char tg[2] = {letter, 0};
  e = Elem();
  e.t = t;
  if (!tag_nc(c, tg) || !two_nodes(c, e)) return false;
  if (!space1(c) || !value_tok(c, e.value)) return false;
  if (t == E_C) {  // capacitor.rs:172 optional G2 flag
    Cur d = c;
    if (space1(d) && tag_nc(d, "G2")) e.g2 = true;
  }
  return true;
}

// "E<name> n+ n- nc+ nc- gain", "G<name> n+ n- nc+ nc- gm",
// "F<name> n+ n- <source> gain", "H<name> n+ n- <source> r" (SPICE syntax)
bool parse_controlled_line(const std::string& line, char letter, ET t, Elem& e) {
  Cur c{line, 0};
  char tg[2] = {letter, 0};
  e = Elem();
  e.t = t;
  if (!tag_nc(c, tg) || !two_nodes(c, e) || !space1(c)) return false;
  if (t == E_E || t == E_G) {
    if (!ident1(c, e.cplus) || !space1(c) || !ident1(c, e.cminus) || !space1(c)) return false;
  } else if (!ident1(c, e.ctrl) || !space1(c)) {
    return false;
  }
  return value_tok(c, e.value);
}

bool parse_diode_line(const std::string& line, Elem& e) {  // diode.rs:75-93
  Cur c{line, 0};
  e = Elem();
  e.t = E_D;
  if (!tag_nc(c, "D") || !two_nodes(c, e)) return false;
  Cur d = c;
  std::string m;
  if (space1(d) && ident1(d, m)) e.model = m;
  else e.model = "default";
  e.value = 1e-12;
  return true;
}

bool parse_wave(Cur& c, const char* tag, int nvals, double* out) {
  Cur d = c;
  if (!tag_nc(d, tag)) return false;
  space0(d);
  if (!tag_exact(d, "(")) return false;
  for (int k = 0; k < nvals; k++) {
    if (k == 0) space0(d);
    else if (!space1(d)) return false;
    if (!value_tok(d, out[k])) return false;
  }
  space0(d);
  if (!tag_exact(d, ")")) return false;
  c = d;
  return true;
}

// voltage_source.rs:187-235
bool parse_vsource(const std::string& line, Elem& e, bool& hard_fail) {
  Cur c{line, 0};
  e = Elem();
  e.t = E_V;
  hard_fail = false;
  if (!tag_nc(c, "V") || !two_nodes(c, e)) return false;
  double dc = 0.0;
  {
    Cur d = c;
    double v;
    if (space1(d)) {
      int r = nom_double(d, v);
      if (r < 0) { hard_fail = true; return false; }
      if (r == 1) { dc = v; c = d; }
    }
  }
  for (;;) {
    Cur d = c;
    if (!space1(d)) break;
    Cur a = d;
    double v;
    if (tag_nc(a, "dc") && space1(a)) {
      int r = nom_double(a, v);
      if (r < 0) { hard_fail = true; return false; }
      if (r == 1) { dc = v; c = a; continue; }
    }
    a = d;
    if (tag_nc(a, "ac") && space1(a)) {
      int r = nom_double(a, v);
      if (r < 0) { hard_fail = true; return false; }
      if (r == 1) { e.ac = v; c = a; continue; }
    }
    a = d;
    double p[7];
    if (parse_wave(a, "pulse", 7, p)) {
      e.w.kind = 1;
      std::copy(p, p + 7, e.w.p);
      c = a;
      continue;
    }
    a = d;
    if (parse_wave(a, "sin", 6, p)) {
      e.w.kind = 2;
      std::fill(e.w.p, e.w.p + 7, 0.0);
      std::copy(p, p + 6, e.w.p);
      c = a;
      continue;
    }
    break;
  }
  e.value = dc;
  return true;
}

// Pulse::value_at voltage_source.rs:68-93; Sinusoidal::value_at 50-65
double wave_value(const Elem& e, double t) {
  const double* p = e.w.p;
  if (e.w.kind == 1) {
    double v1 = p[0], v2 = p[1], td = p[2], tr = p[3], tf = p[4], pw = p[5], per = p[6];
    if (t < td) return v1;
    double tt = std::fmod(t - td, per);
    if (tt < tr) return v1 + (v2 - v1) * (tt / tr);
    if (tt < tr + pw) return v2;
    if (tt < tr + pw + tf) return v2 + (v1 - v2) * ((tt - tr - pw) / tf);
    return v1;
  }
  if (e.w.kind == 2) {
    double vo = p[0], va = p[1], f = p[2], td = p[3], th = p[4], ph = p[5];
    if (t < td) return vo;
    double w = 2.0 * PI_ * f;
    double phr = ph * (PI_ / 180.0);
    double tt = t - td;
    return vo + va * std::exp(-th * tt) * std::sin(w * tt + phr);
  }
  return e.value;  // transient_value_at falls back to dc_value
}

struct DModel {
  double is = 1e-12, rs = 0.0, n = 1.0;
  double cjo = 0.0, vj = 1.0, m = 0.5, fc = 0.5, tt = 0.0;  // junction charge (SPICE), opt-in only
  bool diode = false;
};

// models.rs:40-95 (.model NAME D (k=v ...))
bool parse_model_line(const std::string& raw, std::string& name, DModel& m) {
  std::string line = raw.substr(0, raw.find('%'));
  size_t a = line.find_first_not_of(" \t\r\n"), b = line.find_last_not_of(" \t\r\n");
  line = (a == std::string::npos) ? "" : line.substr(a, b - a + 1);
  Cur c{line, 0};
  if (!tag_nc(c, ".model") || !space1(c) || !ident1(c, name) || !space1(c)) return false;
  bool diode = false;
  if (tag_exact(c, "NMOS") || tag_exact(c, "PMOS")) diode = false;
  else if (tag_exact(c, "D")) diode = true;
  else return false;
  space0(c);
  tag_exact(c, "(");
  std::vector<std::pair<std::string, double>> kv;
  for (;;) {
    Cur d = c;
    space0(d);
    std::string k;
    if (!ident1(d, k)) break;
    space0(d);
    if (!tag_exact(d, "=")) break;
    space0(d);
    double v;
    if (!value_tok(d, v)) break;
    kv.push_back({k, v});
    c = d;
  }
  m = DModel();
  m.diode = diode;
  for (auto& p : kv) {
    std::string k;
    for (char ch : p.first) k += (char)std::tolower((unsigned char)ch);
    if (k == "is") m.is = p.second;
    else if (k == "rs") m.rs = p.second;
    else if (k == "n") m.n = p.second;
    else if (k == "cjo" || k == "cj0") m.cjo = p.second;
    else if (k == "vj") m.vj = p.second;
    else if (k == "m") m.m = p.second;
    else if (k == "fc") m.fc = p.second;
    else if (k == "tt") m.tt = p.second;
  }
  return true;
}

std::string lower(const std::string& s) {
  std::string r;
  for (char ch : s) r += (char)std::tolower((unsigned char)ch);
  return r;
}

// ---- subcircuits (elements/subcircuit.rs:33-200): definitions are collected
// in a first pass over the whole text; an instance "X<name> node... DEF"
// expands in place, every element of the definition mapped: pin names to the
// instance's nodes, any other node (ground "0" too) to "<name>_<node>", the
// element name to "<name>_<element name>"; nested instances the same way.
struct XInst {
  std::string iname, def;
  std::vector<std::string> nodes;
};
struct SubItem {
  bool is_x = false;
  Elem el;
  XInst x;
};
struct SubDef {
  std::vector<std::string> pins;
  std::vector<SubItem> items;
};

// space1 then [A-Za-z0-9_]+ starting at i (utils.rs:67-69); i advances on success
bool take_ident(const std::string& s, size_t& i, std::string& out, bool need_space) {
  size_t j = i;
  if (need_space) {
    size_t b = j;
    while (j < s.size() && (s[j] == ' ' || s[j] == '\t')) j++;
    if (j == b) return false;
  }
  size_t b = j;
  while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) j++;
  if (j == b) return false;
  out = s.substr(b, j - b);
  i = j;
  return true;
}

bool parse_subckt_header(const std::string& line, std::string& name, std::vector<std::string>& pins) {
  if (line.size() < 7 || lower(line.substr(0, 7)) != ".subckt") return false;
  size_t i = 7;
  if (!take_ident(line, i, name, true)) return false;
  std::string p;
  while (take_ident(line, i, p, true)) pins.push_back(p);
  return true;
}

bool parse_x_instance(const std::string& line, XInst& x) {
  if (line.empty() || (line[0] != 'x' && line[0] != 'X')) return false;
  size_t i = 1;
  if (!take_ident(line, i, x.iname, false)) return false;
  std::vector<std::string> toks;
  std::string t;
  while (take_ident(line, i, t, true)) toks.push_back(t);
  if (toks.empty()) return false;
  x.def = toks.back();
  toks.pop_back();
  x.nodes = toks;
  return true;
}

bool parse_primitive(const std::string& line, Elem& el) {
  bool hard = false;
  if (parse_rlci(line, 'R', E_R, el)) return true;
  if (parse_rlci(line, 'C', E_C, el)) return true;
  if (parse_rlci(line, 'L', E_L, el)) return true;
  if (parse_vsource(line, el, hard)) return true;
  if (!hard && parse_rlci(line, 'I', E_I, el)) return true;
  if (hard) return false;
  if (parse_diode_line(line, el)) return true;
  return parse_controlled_line(line, 'E', E_E, el) || parse_controlled_line(line, 'G', E_G, el) ||
         parse_controlled_line(line, 'F', E_F, el) || parse_controlled_line(line, 'H', E_H, el);
}

void expand_instance(const XInst& x, const std::map<std::string, SubDef>& defs, std::vector<Elem>& out, int depth) {
  auto d = defs.find(x.def);
  if (d == defs.end()) throw RefError{REF_PARSE, "undefined subcircuit definition: " + x.def};
  if (x.nodes.size() != d->second.pins.size())
    throw RefError{REF_PARSE, "node/port mismatch for instance " + x.iname};
  if (depth > 64) throw RefError{REF_PARSE, "subcircuit nesting too deep: " + x.def};
  std::map<std::string, std::string> p2n;
  for (size_t k = 0; k < x.nodes.size(); k++) p2n[d->second.pins[k]] = x.nodes[k];
  auto map = [&](const std::string& nd) {
    auto f = p2n.find(nd);
    return f != p2n.end() ? f->second : x.iname + "_" + nd;
  };
  for (const SubItem& it : d->second.items) {
    if (!it.is_x) {
      Elem e = it.el;
      e.plus = map(e.plus);
      e.minus = map(e.minus);
      if (e.t == E_E || e.t == E_G) {
        e.cplus = map(e.cplus);
        e.cminus = map(e.cminus);
      }
      // a controlling element named inside a definition is the instance's own
      if (!e.ctrl.empty()) e.ctrl = e.ctrl.substr(0, 1) + x.iname + "_" + e.ctrl.substr(1);
      e.name = x.iname + "_" + e.name;
      out.push_back(e);
    } else {
      XInst y = it.x;
      for (auto& nd : y.nodes) nd = map(nd);
      y.iname = x.iname + "_" + y.iname;
      expand_instance(y, defs, out, depth + 1);
    }
  }
}

std::map<std::string, SubDef> collect_subcircuits(const std::string& text) {
  std::map<std::string, SubDef> defs;
  std::string cur;
  bool in_def = false;
  size_t pos = 0;
  while (pos < text.size()) {
    size_t e = text.find('\n', pos);
    if (e == std::string::npos) e = text.size();
    std::string line = text.substr(pos, e - pos);
    pos = e + 1;
    size_t a = line.find_first_not_of(" \t\r\n\v\f");
    size_t b = line.find_last_not_of(" \t\r\n\v\f");
    line = (a == std::string::npos) ? "" : line.substr(a, b - a + 1);
    if (line.empty() || line[0] == '%' || line[0] == '*') continue;
    std::string lo = lower(line);
    if (lo.rfind(".subckt", 0) == 0) {
      std::string nm;
      std::vector<std::string> pins;
      if (!parse_subckt_header(line, nm, pins)) throw RefError{REF_PARSE, "bad subcircuit header: " + line};
      defs[nm] = SubDef{pins, {}};
      cur = nm;
      in_def = true;
      continue;
    }
    if (lo.rfind(".ends", 0) == 0) {
      in_def = false;
      continue;
    }
    if (!in_def) continue;
    SubItem it;
    if (!parse_primitive(line, it.el)) {
      if (!parse_x_instance(line, it.x)) {
        char f = (char)std::toupper((unsigned char)line[0]);
        throw RefError{f == 'Q' || f == 'M' ? REF_INVALID_ELEMENT_FORMAT : REF_PARSE,
                       "bad element in subcircuit " + cur + ": " + line};
      }
      it.is_x = true;
    }
    defs[cur].items.push_back(it);
  }
  return defs;
}

// flags & 1: the opt-in diode parasitics (series resistance as an internal
// node and resistor numbered after krets' own, junction charge parameters).
Circuit* parse_netlist(const std::string& text, int flags = 0) {
  const std::map<std::string, SubDef> defs = collect_subcircuits(text);
  auto* C = new Circuit();
  std::map<std::string, DModel> models;
  bool in_ctl = false, in_sub = false;
  size_t pos = 0;
  int lineno = 0;
  while (pos <= text.size()) {
    size_t e = text.find('\n', pos);
    if (e == std::string::npos) e = text.size();
    std::string line = text.substr(pos, e - pos);
    pos = e + 1;
    lineno++;
    size_t a = line.find_first_not_of(" \t\r\n\v\f");
    size_t b = line.find_last_not_of(" \t\r\n\v\f");
    line = (a == std::string::npos) ? "" : line.substr(a, b - a + 1);
    if (e == text.size() && line.empty() && pos > text.size()) break;
    if (line.empty() || line[0] == '%' || line[0] == '*') continue;
    std::string lo = lower(line);
    if (lo.rfind(".control", 0) == 0) { in_ctl = true; continue; }
    if (lo.rfind(".endc", 0) == 0) { in_ctl = false; continue; }
    if (in_ctl) continue;
    if (lo.rfind(".subckt", 0) == 0) { in_sub = true; continue; }
    if (lo.rfind(".ends", 0) == 0) { in_sub = false; continue; }
    if (in_sub) continue;
    if (lo == ".end") continue;
    if (lo.rfind(".model", 0) == 0) {
      std::string nm;
      DModel m;
      if (!parse_model_line(line, nm, m)) {
        delete C;
        throw RefError{REF_PARSE, "line " + std::to_string(lineno) + ": bad .model"};
      }
      models[nm] = m;
      continue;
    }
    Elem el;
    if (!parse_primitive(line, el)) {
      XInst x;
      if (parse_x_instance(line, x)) {
        try {
          expand_instance(x, defs, C->el, 0);
        } catch (...) {
          delete C;
          throw;
        }
        continue;
      }
      char f = (char)std::toupper((unsigned char)line[0]);
      delete C;
      if (f == 'Q' || f == 'M')
        throw RefError{REF_INVALID_ELEMENT_FORMAT,
                       "line " + std::to_string(lineno) + ": element outside the R/L/C/V/I/D subset"};
      throw RefError{REF_PARSE, "line " + std::to_string(lineno) + ": cannot parse '" + line + "'"};
    }
    C->el.push_back(el);
  }
  // parser.rs:105-125 index map
  std::set<std::string> seen;
  int counter = 0;
  for (auto& el : C->el) {
    if (el.is_g2()) {
      std::string k = "I(" + el.ident() + ")";
      C->index[k] = counter++;
    }
    std::vector<const std::string*> nodes{&el.plus, &el.minus};
    if (el.t == E_E || el.t == E_G) nodes.insert(nodes.end(), {&el.cplus, &el.cminus});
    for (const std::string* nd : nodes) {
      if (seen.insert(*nd).second) {
        if (*nd == "0") continue;
        C->index["V(" + *nd + ")"] = counter++;
      }
    }
  }
  if (C->el.empty()) {
    delete C;
    throw RefError{REF_EMPTY_NETLIST, "empty netlist"};
  }
  if ((int)C->index.size() != counter) {
    delete C;
    throw RefError{REF_UNEXPECTED, "duplicate element identifiers"};
  }
  C->names.assign(counter, "");
  for (auto& kv : C->index) C->names[kv.second] = kv.first;
  // parser.rs:131-149 model binding (missing model panics in krets)
  for (auto& el : C->el) {
    if (el.t == E_D) {
      auto it = models.find(el.model);
      if (it == models.end() || !it->second.diode) {
        std::string msg = "diode " + el.ident() + " has no .model '" + el.model + "'";
        delete C;
        throw RefError{REF_UNEXPECTED, msg};
      }
      el.value = it->second.is;
      el.n_emis = it->second.n;
      if (flags & 1) {
        const DModel& md = it->second;
        el.jq[0] = md.cjo;
        el.jq[1] = md.vj;
        el.jq[2] = md.m;
        el.jq[3] = md.fc;
        el.jq[4] = md.tt;
      }
      C->nonlinear = true;
    }
    auto f = [&](const std::string& nd) {
      auto it = C->index.find("V(" + nd + ")");
      return it == C->index.end() ? -1 : it->second;
    };
    el.ip = f(el.plus);
    el.im = f(el.minus);
    auto it = C->index.find("I(" + el.ident() + ")");
    el.ic = it == C->index.end() ? -1 : it->second;
    if (el.t == E_E || el.t == E_G) {
      el.icp = f(el.cplus);
      el.icm = f(el.cminus);
    }
    if (el.t == E_F || el.t == E_H) {
      std::string id = el.ctrl;
      id[0] = (char)std::toupper((unsigned char)id[0]);
      auto b = C->index.find("I(" + id + ")");
      if (b == C->index.end()) {
        std::string msg = "controlling element " + id + " of " + el.ident() + " not found";
        delete C;
        throw RefError{REF_ELEMENT_NOT_FOUND, msg};
      }
      el.ctrl = id;
      el.icb = b->second;
    }
  }
  if (flags & 1) {
    // series resistance: an internal anode node V(<id>:a) and a resistor
    // <id>:rs from the anode to it, after every krets unknown and element
    const size_t ne0 = C->el.size();
    for (size_t k = 0; k < ne0; k++) {
      Elem& d = C->el[k];
      if (d.t != E_D) continue;
      const double rs = models.at(d.model).rs;
      if (!(rs > 0.0)) continue;
      const std::string inode = d.ident() + ":a";
      const int ia = (int)C->names.size();
      C->names.push_back("V(" + inode + ")");
      C->index["V(" + inode + ")"] = ia;
      Elem r;
      r.t = E_R;
      r.name = d.name + ":rs";
      r.plus = d.plus;
      r.minus = inode;
      r.value = rs;
      r.ip = d.ip;
      r.im = ia;
      d.plus = inode;
      d.ip = ia;
      C->el.push_back(r);
    }
  }
  return C;
}

// ------------------------------------------------------- sparse LU (faer stand-in)
template <class T>
double mag(const T& v) { return std::abs(v); }

// Minimum-degree order on the pattern of A + A^T.
std::vector<int> min_degree(int n, const std::vector<int>& cp, const std::vector<int>& ri) {
  std::vector<std::set<int>> adj(n);
  for (int j = 0; j < n; j++)
    for (int k = cp[j]; k < cp[j + 1]; k++) {
      int i = ri[k];
      if (i != j) {
        adj[i].insert(j);
        adj[j].insert(i);
      }
    }
  std::set<std::pair<int, int>> pq;
  for (int v = 0; v < n; v++) pq.insert({(int)adj[v].size(), v});
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> alive(n, 1);
  while (!pq.empty()) {
    int v = pq.begin()->second;
    pq.erase(pq.begin());
    order.push_back(v);
    alive[v] = 0;
    std::vector<int> nb(adj[v].begin(), adj[v].end());
    for (int u : nb) {
      pq.erase({(int)adj[u].size(), u});
      adj[u].erase(v);
      for (int w : nb)
        if (w != u) adj[u].insert(w);
      pq.insert({(int)adj[u].size(), u});
    }
    adj[v].clear();
  }
  return order;
}

template <class T>
struct SpLU {
  int n = 0;
  std::vector<int> q, pinv, prow;
  std::vector<int> Lp, Li, Up, Ui;
  std::vector<T> Lx, Ux;

  bool factor(int n_, const std::vector<int>& cp, const std::vector<int>& ri, const std::vector<T>& vx,
              const std::vector<int>& order) {
    n = n_;
    q = order;
    pinv.assign(n, -1);
    prow.assign(n, -1);
    Lp.assign(1, 0);
    Up.assign(1, 0);
    Li.clear(); Lx.clear(); Ui.clear(); Ux.clear();
    std::vector<T> x(n, T(0));
    std::vector<char> mark(n, 0), inpat(n, 0);
    std::vector<int> stack, post, pat, iter;
    for (int k = 0; k < n; k++) {
      int col = q[k];
      post.clear();
      pat.clear();
      // DFS reach over pivotal rows through L columns (topological order)
      for (int p = cp[col]; p < cp[col + 1]; p++) {
        int r = ri[p];
        if (!inpat[r]) { inpat[r] = 1; pat.push_back(r); }
        if (pinv[r] < 0 || mark[r]) continue;
        stack.assign(1, r);
        iter.assign(1, Lp[pinv[r]]);
        mark[r] = 1;
        while (!stack.empty()) {
          int v = stack.back();
          int j = pinv[v];
          int& it = iter.back();
          bool pushed = false;
          while (it < Lp[j + 1]) {
            int w = Li[it++];
            if (pinv[w] >= 0 && !mark[w]) {
              mark[w] = 1;
              stack.push_back(w);
              iter.push_back(Lp[pinv[w]]);
              pushed = true;
              break;
            }
          }
          if (!pushed) {
            post.push_back(v);
            stack.pop_back();
            iter.pop_back();
          }
        }
      }
      for (int p = cp[col]; p < cp[col + 1]; p++) x[ri[p]] += vx[p];
      for (int t = (int)post.size() - 1; t >= 0; t--) {
        int r = post[t];
        int j = pinv[r];
        T xr = x[r];
        for (int p = Lp[j]; p < Lp[j + 1]; p++) {
          int i = Li[p];
          if (!inpat[i]) { inpat[i] = 1; pat.push_back(i); }
          x[i] -= Lx[p] * xr;
        }
      }
      int piv = -1;
      double best = 0.0;
      for (int i : pat) {
        if (pinv[i] >= 0) continue;
        double a = mag(x[i]);
        if (a > best || (a == best && a > 0 && i < piv)) { best = a; piv = i; }
      }
      for (int t = (int)post.size() - 1; t >= 0; t--) {
        int r = post[t];
        Ui.push_back(pinv[r]);
        Ux.push_back(x[r]);
      }
      if (piv < 0 || !(best > 0.0) || !std::isfinite(best)) {
        for (int i : pat) { x[i] = T(0); inpat[i] = 0; }
        for (int r : post) mark[r] = 0;
        return false;
      }
      T pv = x[piv];
      Ui.push_back(k);
      Ux.push_back(pv);
      Up.push_back((int)Ui.size());
      for (int i : pat) {
        if (pinv[i] >= 0 || i == piv) continue;
        Li.push_back(i);
        Lx.push_back(x[i] / pv);
      }
      Lp.push_back((int)Li.size());
      pinv[piv] = k;
      prow[k] = piv;
      for (int i : pat) { x[i] = T(0); inpat[i] = 0; }
      for (int r : post) mark[r] = 0;
    }
    return true;
  }

  void solve(const std::vector<T>& b, std::vector<T>& y) const {
    std::vector<T> w(b), z(n);
    std::vector<T> c(n);
    for (int j = 0; j < n; j++) {
      T wj = w[prow[j]];
      c[j] = wj;
      for (int p = Lp[j]; p < Lp[j + 1]; p++) w[Li[p]] -= Lx[p] * wj;
    }
    for (int k = n - 1; k >= 0; k--) {
      T d = Ux[Up[k + 1] - 1];
      T zk = c[k] / d;
      z[k] = zk;
      for (int p = Up[k]; p < Up[k + 1] - 1; p++) c[Ui[p]] -= Ux[p] * zk;
    }
    y.assign(n, T(0));
    for (int k = 0; k < n; k++) y[q[k]] = z[k];
  }

  // A^T y = b
  void solve_t(const std::vector<T>& b, std::vector<T>& y) const {
    std::vector<T> v(n);
    for (int k = 0; k < n; k++) {
      T s = b[q[k]];
      for (int p = Up[k]; p < Up[k + 1] - 1; p++) s -= Ux[p] * v[Ui[p]];
      v[k] = s / Ux[Up[k + 1] - 1];
    }
    y.assign(n, T(0));
    for (int j = n - 1; j >= 0; j--) {
      T s = v[j];
      for (int p = Lp[j]; p < Lp[j + 1]; p++) s -= Lx[p] * y[Li[p]];
      y[prow[j]] = s;
    }
  }
};

template <class T>
struct Trip {
  int r, c;
  T v;
};

// Sum duplicate triplets (solver.rs:59-70) and build CSC.
template <class T>
void sum_to_csc(int n, const std::vector<Trip<T>>& tr, int cost, std::vector<int>& cp, std::vector<int>& ri,
                std::vector<T>& vx) {
  std::vector<Trip<T>> s;
  if (cost == REF_COST_COMPAT) {
    std::unordered_map<uint64_t, T> m;
    for (auto& t : tr) m[(uint64_t)t.r * (uint64_t)n + (uint64_t)t.c] += t.v;
    s.reserve(m.size());
    for (auto& kv : m) s.push_back({(int)(kv.first / n), (int)(kv.first % n), kv.second});
  } else {
    s = tr;
  }
  std::sort(s.begin(), s.end(), [](const Trip<T>& a, const Trip<T>& b) { return a.c != b.c ? a.c < b.c : a.r < b.r; });
  cp.assign(n + 1, 0);
  ri.clear();
  vx.clear();
  for (size_t i = 0; i < s.size();) {
    size_t j = i;
    T v = s[i].v;
    while (++j < s.size() && s[j].r == s[i].r && s[j].c == s[i].c) v += s[j].v;
    ri.push_back(s[i].r);
    vx.push_back(v);
    cp[s[i].c + 1]++;
    i = j;
  }
  for (int j = 0; j < n; j++) cp[j + 1] += cp[j];
}

template <class T>
struct LinSolver {
  int cost = REF_COST_COMPAT;
  std::vector<int> cached_order, cached_cp, cached_ri;
  SpLU<T> lu;
  bool solve(int n, const std::vector<Trip<T>>& A, const std::vector<Trip<T>>& b, std::vector<T>& x) {
    std::vector<int> cp, ri;
    std::vector<T> vx;
    sum_to_csc(n, A, cost, cp, ri, vx);
    std::vector<int> order;
    if (cost == REF_COST_LEAN && cp == cached_cp && ri == cached_ri) {
      order = cached_order;
    } else {
      order = min_degree(n, cp, ri);
      if (cost == REF_COST_LEAN) { cached_order = order; cached_cp = cp; cached_ri = ri; }
    }
    if (!lu.factor(n, cp, ri, vx, order)) return false;
    std::vector<T> rhs(n, T(0));
    if (cost == REF_COST_COMPAT) {
      std::unordered_map<uint64_t, T> m;
      for (auto& t : b) m[(uint64_t)t.r] += t.v;
      for (auto& kv : m) rhs[kv.first] = kv.second;
    } else {
      for (auto& t : b) rhs[t.r] += t.v;
    }
    lu.solve(rhs, x);
    return true;
  }
};

// ----------------------------------------------------------- solution maps
// COMPAT keeps krets' HashMap<String, f64>; lookups format key strings.
struct SolMap {
  bool compat;
  bool empty = true;
  std::unordered_map<std::string, double> m;  // compat
  std::vector<double> x;                      // lean
  double time = 0.0;
  bool has_time = false;
  const Circuit* C;
  SolMap(const Circuit* c, bool cm) : compat(cm), C(c) {}
  bool getV(const Elem& e, bool plus, double& v) const {
    if (compat) {
      auto it = m.find("V(" + (plus ? e.plus : e.minus) + ")");
      if (it == m.end()) return false;
      v = it->second;
      return true;
    }
    if (empty) return false;
    int i = plus ? e.ip : e.im;
    if (i < 0) return false;
    v = x[i];
    return true;
  }
  bool getI(const Elem& e, double& v) const {
    if (compat) {
      auto it = m.find("I(" + e.ident() + ")");
      if (it == m.end()) return false;
      v = it->second;
      return true;
    }
    if (empty || e.ic < 0) return false;
    v = x[e.ic];
    return true;
  }
  double get_time() const {
    if (compat) {
      auto it = m.find("time");
      return it == m.end() ? 0.0 : it->second;
    }
    return has_time ? time : 0.0;
  }
  // op.rs:60-63 result map build
  void set(const std::vector<double>& sol, const char* extra_key, double extra) {
    empty = false;
    if (compat) {
      m.clear();
      for (auto& kv : C->index) m[kv.first] = sol[kv.second];
      if (extra_key) m[extra_key] = extra;
    } else {
      x = sol;
    }
    has_time = extra_key && std::strcmp(extra_key, "time") == 0;
    if (has_time) time = extra;
    extra_val = extra;
    has_extra = extra_key != nullptr;
  }
  double extra_val = 0.0;
  bool has_extra = false;
};

// solver.rs:76-104 convergence_check
bool converged(const SolMap& prev, const SolMap& cur, const ref_config& cfg) {
  if (prev.empty) return false;
  auto ok = [&](double v, double pv, bool cur_tol) {
    double diff = std::fabs(v - pv);
    double scale = std::max(std::fabs(v), std::fabs(pv));
    double atol = cur_tol ? cfg.iabstol : cfg.vabstol;
    return diff <= cfg.reltol * scale + atol;
  };
  if (cur.compat) {
    for (auto& kv : cur.m) {
      auto it = prev.m.find(kv.first);
      double pv = it == prev.m.end() ? 0.0 : it->second;
      if (!ok(kv.second, pv, kv.first[0] == 'I')) return false;
    }
    return true;
  }
  for (size_t i = 0; i < cur.x.size(); i++)
    if (!ok(cur.x[i], prev.x[i], cur.C->names[i][0] == 'I')) return false;
  if (cur.has_extra) {
    double pv = prev.has_extra ? prev.extra_val : 0.0;
    if (!ok(cur.extra_val, pv, false)) return false;
  }
  return true;
}

// ------------------------------------------------------------- stamps
struct Idx {
  const Circuit* C;
  bool compat;
  // 0 plus, 1 minus, 2 current, 3 / 4 controlling plus / minus, 5 controlling current
  int get(const Elem& e, int which) const {
    if (!compat) {
      const int v[] = {e.ip, e.im, e.ic, e.icp, e.icm, e.icb};
      return v[which];
    }
    if ((which == 3 || which == 4) && e.t != E_E && e.t != E_G) return -1;
    if (which == 5 && e.t != E_F && e.t != E_H) return -1;
    std::string k;
    switch (which) {
      case 0: k = "V(" + e.plus + ")"; break;
      case 1: k = "V(" + e.minus + ")"; break;
      case 2: k = "I(" + e.ident() + ")"; break;
      case 3: k = "V(" + e.cplus + ")"; break;
      case 4: k = "V(" + e.cminus + ")"; break;
      default: k = "I(" + e.ctrl + ")"; break;
    }
    auto it = C->index.find(k);
    return it == C->index.end() ? -1 : it->second;
  }
};

// Linear controlled sources, the same in every analysis (SPICE conventions):
//   E: branch row V(p) - V(m) - k (V(cp) - V(cm)) = 0, branch current at p, m
//   G: k (V(cp) - V(cm)) flows from p through the source to m
//   F: k I(cb) flows from p through the source to m
//   H: branch row V(p) - V(m) - k I(cb) = 0, branch current at p, m
template <class T>
void stamp_controlled(const Elem& e, const Idx& ix, std::vector<Trip<T>>& A) {
  int p = ix.get(e, 0), m = ix.get(e, 1), c = ix.get(e, 2);
  int cp = ix.get(e, 3), cm = ix.get(e, 4), cb = ix.get(e, 5);
  const T k(e.value), one(1.0);
  auto put = [&](int r, int col, T v) {
    if (r >= 0 && col >= 0) A.push_back({r, col, v});
  };
  if (e.t == E_E || e.t == E_H) {
    put(p, c, one);
    put(c, p, one);
    put(m, c, -one);
    put(c, m, -one);
  }
  switch (e.t) {
    case E_E: put(c, cp, -k); put(c, cm, k); break;
    case E_G: put(p, cp, k); put(p, cm, -k); put(m, cp, -k); put(m, cm, k); break;
    case E_F: put(p, cb, k); put(m, cb, -k); break;
    case E_H: put(c, cb, -k); break;
    default: break;
  }
}

template <class T>
void four(std::vector<Trip<T>>& A, int p, int m, T y) {
  if (p >= 0) A.push_back({p, p, y});
  if (m >= 0) A.push_back({m, m, y});
  if (p >= 0 && m >= 0) {
    A.push_back({p, m, -y});
    A.push_back({m, p, -y});
  }
}

// diode.rs:24-71 (krets: limit is +-inf for any practical Is)
struct DiodeLin {
  double g, ieq;
};
DiodeLin diode_compat(const Elem& e, const SolMap& s) {
  double vp = 0.5, vm = 0.0, t;
  if (s.getV(e, true, t)) vp = t;
  if (s.getV(e, false, t)) vm = t;
  double vd = vp - vm;
  double n = e.n_emis, is = e.value;
  double vcrit = n * VT * std::log(1.7976931348623157e308 * n * VT / is);
  double vl = std::min(std::max(vd, -vcrit), vcrit);
  double g = (is / (n * VT)) * std::exp(vl / (n * VT));
  double i = is * (std::exp(vl / (n * VT)) - 1.0);
  return {g, i - g * vd};
}

// stampable.rs DC stamps (matrix + excitation)
void stamp_dc(const Circuit& C, const Idx& ix, const SolMap& sol, std::vector<Trip<double>>& A,
              std::vector<Trip<double>>& b, bool skip_caps, int mode_tran, double h, const SolMap* prev) {
  for (const Elem& e : C.el) {
    if (skip_caps && e.t == E_C) continue;
    int p = ix.get(e, 0), m = ix.get(e, 1);
    switch (e.t) {
      case E_R: {
        if (e.g2) {
          int c = ix.get(e, 2);
          if (p >= 0 && c >= 0) { A.push_back({p, c, 1.0}); A.push_back({c, p, 1.0}); }
          if (m >= 0 && c >= 0) { A.push_back({m, c, -1.0}); A.push_back({c, m, -1.0}); }
          if (c >= 0) A.push_back({c, c, -e.value});
        } else {
          four(A, p, m, 1.0 / e.value);
        }
        break;
      }
      case E_C: {
        if (!mode_tran) break;
        double g = e.value / h;
        four(A, p, m, g);
        double vp = 0, vm = 0, t;
        if (prev->getV(e, true, t)) vp = t;
        if (prev->getV(e, false, t)) vm = t;
        double ieq = -(e.value / h) * (vp - vm);
        if (p >= 0) b.push_back({p, 0, -ieq});
        if (m >= 0) b.push_back({m, 0, ieq});
        break;
      }
      case E_L: {
        int c = ix.get(e, 2);
        if (mode_tran && c >= 0) A.push_back({c, c, -e.value / h});
        if (p >= 0 && c >= 0) { A.push_back({p, c, 1.0}); A.push_back({c, p, 1.0}); }
        if (m >= 0 && c >= 0) { A.push_back({m, c, -1.0}); A.push_back({c, m, -1.0}); }
        if (mode_tran) {
          double ip;
          if (!prev->getI(e, ip)) throw RefError{REF_UNEXPECTED, "inductor without previous current"};
          if (c >= 0) b.push_back({c, 0, -(e.value / h) * ip});
        }
        break;
      }
      case E_V: {
        int c = ix.get(e, 2);
        if (p >= 0 && c >= 0) { A.push_back({p, c, 1.0}); A.push_back({c, p, 1.0}); }
        if (m >= 0 && c >= 0) { A.push_back({m, c, -1.0}); A.push_back({c, m, -1.0}); }
        if (c >= 0) b.push_back({c, 0, mode_tran ? wave_value(e, sol.get_time()) : e.value});
        break;
      }
      case E_I: {
        int c = ix.get(e, 2);
        if (c >= 0) A.push_back({c, c, 1.0});
        if (p >= 0 && c >= 0) A.push_back({p, c, 1.0});
        if (m >= 0 && c >= 0) A.push_back({m, c, -1.0});
        if (c >= 0) b.push_back({c, 0, e.value});
        break;
      }
      case E_D: {
        DiodeLin d = diode_compat(e, sol);
        four(A, p, m, d.g);
        if (p >= 0) b.push_back({p, 0, -d.ieq});
        if (m >= 0) b.push_back({m, 0, d.ieq});
        break;
      }
      default: stamp_controlled(e, ix, A); break;
    }
  }
}

std::pair<double, double> junction_charge(const Elem& e, double v);

void stamp_ac(const Circuit& C, const Idx& ix, const SolMap& op, double f, std::vector<Trip<cd>>& A,
              std::vector<Trip<cd>>& b) {
  for (const Elem& e : C.el) {
    int p = ix.get(e, 0), m = ix.get(e, 1);
    switch (e.t) {
      case E_R: {
        if (e.g2) {
          int c = ix.get(e, 2);
          cd one(1.0, 0.0);
          if (p >= 0 && c >= 0) { A.push_back({p, c, one}); A.push_back({c, p, one}); }
          if (m >= 0 && c >= 0) { A.push_back({m, c, -one}); A.push_back({c, m, -one}); }
          if (c >= 0) A.push_back({c, c, -cd(e.value, 0.0)});
        } else {
          four(A, p, m, cd(1.0 / e.value, 0.0));
        }
        break;
      }
      case E_C: {
        cd y(0.0, 2.0 * PI_ * f * e.value);
        if (!e.g2) {
          four(A, p, m, y);
        } else {
          int c = ix.get(e, 2);
          if (p >= 0 && c >= 0) A.push_back({c, p, -y});
          if (m >= 0 && c >= 0) A.push_back({c, m, y});
          if (c >= 0) A.push_back({c, c, cd(1.0, 0.0)});
        }
        break;
      }
      case E_L: {
        int c = ix.get(e, 2);
        cd z(0.0, 2.0 * PI_ * f * e.value);
        if (p >= 0 && c >= 0) { A.push_back({p, c, cd(1, 0)}); }
        if (m >= 0 && c >= 0) { A.push_back({m, c, cd(-1, 0)}); }
        if (p >= 0 && c >= 0) { A.push_back({c, p, cd(1, 0)}); }
        if (m >= 0 && c >= 0) { A.push_back({c, m, cd(-1, 0)}); }
        if (c >= 0) A.push_back({c, c, -z});
        break;
      }
      case E_V: {
        int c = ix.get(e, 2);
        cd one(1.0, 0.0);
        if (p >= 0 && c >= 0) { A.push_back({p, c, one}); A.push_back({c, p, one}); }
        if (m >= 0 && c >= 0) { A.push_back({m, c, -one}); A.push_back({c, m, -one}); }
        if (c >= 0) b.push_back({c, 0, cd(e.ac, 0.0)});
        break;
      }
      case E_I: {
        int c = ix.get(e, 2);
        if (p >= 0 && c >= 0) A.push_back({p, c, cd(1, 0)});
        if (m >= 0 && c >= 0) A.push_back({m, c, cd(-1, 0)});
        if (c >= 0) b.push_back({c, 0, cd(e.value, 0.0)});
        break;
      }
      case E_D: {
        DiodeLin d = diode_compat(e, op);
        double c = 0.0;
        if (e.has_jq()) {  // junction capacitance at the operating point (opt-in parasitics)
          double vp = 0.5, vm = 0.0, tv;
          if (op.getV(e, true, tv)) vp = tv;
          if (op.getV(e, false, tv)) vm = tv;
          c = junction_charge(e, vp - vm).second;
        }
        four(A, p, m, cd(d.g, 2.0 * PI_ * f * c));
        break;
      }
      default: stamp_controlled(e, ix, A); break;
    }
  }
}

int N_of(const Circuit& C) { return (int)C.names.size(); }

// solver/op.rs:15-86
std::vector<double> op_solve(const Circuit& C, const ref_config& cfg, int* iters_out) {
  int n = N_of(C);
  bool compat = cfg.cost == REF_COST_COMPAT;
  Idx ix{&C, compat};
  LinSolver<double> ls;
  ls.cost = cfg.cost;
  SolMap prev(&C, compat), res(&C, compat);
  std::vector<double> x;
  for (int it = 0; it < cfg.maxiter; it++) {
    std::vector<Trip<double>> A, b;
    stamp_dc(C, ix, prev, A, b, true, 0, 0.0, nullptr);
    if (!ls.solve(n, A, b, x)) throw RefError{REF_DECOMPOSITION, "OP: matrix decomposition failed"};
    res.set(x, nullptr, 0.0);
    if (iters_out) *iters_out = it + 1;
    if (!C.nonlinear) break;
    if (converged(prev, res, cfg)) break;
    prev = res;
    if (it == cfg.maxiter - 1) throw RefError{REF_MAX_ITERATIONS, "OP: maximum iterations exceeded"};
  }
  return x;
}

// --------------------------------------------------- clean-mode helpers
// Diode junction charge (opt-in parasitics; SPICE2 diode model, Nagel 1975):
// depletion charge CJO VJ / (1 - M) (1 - (1 - v/VJ)^(1-M)) and capacitance
// CJO (1 - v/VJ)^-M below FC VJ, continued linearly in the capacitance above
// it, plus the diffusion charge TT * I(v). Returns (q, c) at v.
std::pair<double, double> junction_charge(const Elem& e, double v) {
  const double cjo = e.jq[0], vj = e.jq[1], m = e.jq[2], fc = e.jq[3], tt = e.jq[4];
  const double nvt = e.n_emis * VT, ex = std::exp(v / nvt);
  double q = tt * e.value * (ex - 1.0), c = tt * e.value / nvt * ex;
  if (cjo != 0.0) {
    if (v < fc * vj) {
      q += cjo * vj / (1.0 - m) * (1.0 - std::pow(1.0 - v / vj, 1.0 - m));
      c += cjo * std::pow(1.0 - v / vj, -m);
    } else {
      const double f1 = vj / (1.0 - m) * (1.0 - std::pow(1.0 - fc, 1.0 - m));
      const double f2 = std::pow(1.0 - fc, 1.0 + m), f3 = 1.0 - fc * (1.0 + m);
      const double v0 = fc * vj;
      q += cjo * (f1 + (f3 * (v - v0) + m / (2.0 * vj) * (v * v - v0 * v0)) / f2);
      c += cjo / f2 * (f3 + m * v / vj);
    }
  }
  return {q, c};
}

// SPICE3 pnjlim junction-voltage limiting.
double pnjlim(double vnew, double vold, double vt, double vcrit, bool& limited) {
  limited = false;
  if (vnew > vcrit && std::fabs(vnew - vold) > vt + vt) {
    if (vold > 0) {
      double arg = 1.0 + (vnew - vold) / vt;
      vnew = arg > 0 ? vold + vt * std::log(arg) : vcrit;
    } else {
      vnew = vt * std::log(vnew / vt);
    }
    limited = true;
  }
  return vnew;
}

struct CleanState {
  std::vector<double> ic;  // capacitor currents (TRAP)
  std::vector<double> vl;  // inductor voltages (TRAP)
  std::vector<double> iq;  // diode junction-charge currents (TRAP)
};

double vdiff(const Elem& e, const std::vector<double>& x) {
  double vp = e.ip >= 0 ? x[e.ip] : 0.0, vm = e.im >= 0 ? x[e.im] : 0.0;
  return vp - vm;
}

// One clean Newton solve. kind: 0 = OP, 1 = BE step, 2 = TRAP step.
// Returns iteration count; throws on failure.
int clean_newton(const Circuit& C, const ref_config& cfg, int kind, double t, double h,
                 const std::vector<double>& xprev, const CleanState& st, std::vector<double>& x) {
  int n = N_of(C);
  LinSolver<double> ls;
  ls.cost = REF_COST_LEAN;
  std::vector<double> guess = xprev;
  std::vector<double> vdold(C.el.size(), 0.0);
  for (size_t k = 0; k < C.el.size(); k++)
    if (C.el[k].t == E_D) vdold[k] = vdiff(C.el[k], guess);
  double alpha = kind == 2 ? 2.0 : 1.0;
  const Idx lean{&C, false};
  for (int it = 0; it < cfg.maxiter; it++) {
    std::vector<Trip<double>> A, b;
    bool any_lim = false;
    for (size_t k = 0; k < C.el.size(); k++) {
      const Elem& e = C.el[k];
      int p = e.ip, m = e.im, c = e.ic;
      switch (e.t) {
        case E_R:
          if (e.g2) {
            if (p >= 0 && c >= 0) { A.push_back({p, c, 1.0}); A.push_back({c, p, 1.0}); }
            if (m >= 0 && c >= 0) { A.push_back({m, c, -1.0}); A.push_back({c, m, -1.0}); }
            if (c >= 0) A.push_back({c, c, -e.value});
          } else {
            four(A, p, m, 1.0 / e.value);
          }
          break;
        case E_C: {
          if (kind == 0) break;
          double g = alpha * e.value / h;
          four(A, p, m, g);
          double ieq = g * vdiff(e, xprev) + (kind == 2 ? st.ic[k] : 0.0);
          if (p >= 0) b.push_back({p, 0, ieq});
          if (m >= 0) b.push_back({m, 0, -ieq});
          break;
        }
        case E_L: {
          if (p >= 0 && c >= 0) { A.push_back({p, c, 1.0}); A.push_back({c, p, 1.0}); }
          if (m >= 0 && c >= 0) { A.push_back({m, c, -1.0}); A.push_back({c, m, -1.0}); }
          if (kind != 0) {
            double z = alpha * e.value / h;
            A.push_back({c, c, -z});
            b.push_back({c, 0, -z * xprev[c] - (kind == 2 ? st.vl[k] : 0.0)});
          }
          break;
        }
        case E_V:
          if (p >= 0) { A.push_back({p, c, 1.0}); A.push_back({c, p, 1.0}); }
          if (m >= 0) { A.push_back({m, c, -1.0}); A.push_back({c, m, -1.0}); }
          b.push_back({c, 0, kind == 0 ? e.value : wave_value(e, t)});
          break;
        case E_I:
          A.push_back({c, c, 1.0});
          if (p >= 0) A.push_back({p, c, 1.0});
          if (m >= 0) A.push_back({m, c, -1.0});
          b.push_back({c, 0, e.value});
          break;
        case E_D: {
          double nvt = e.n_emis * VT;
          double vcrit = nvt * std::log(nvt / (std::sqrt(2.0) * e.value));
          bool lim;
          double vd = pnjlim(vdiff(e, guess), vdold[k], nvt, vcrit, lim);
          any_lim |= lim;
          vdold[k] = vd;
          double ex = std::exp(vd / nvt);
          double g = e.value / nvt * ex;
          double id = e.value * (ex - 1.0);
          if (kind != 0 && e.has_jq()) {  // junction-charge companion of the BE / TRAP step
            const auto q = junction_charge(e, vd), qo = junction_charge(e, vdiff(e, xprev));
            const double gq = alpha * q.second / h;
            id += alpha * (q.first - qo.first) / h - (alpha - 1.0) * (st.iq.empty() ? 0.0 : st.iq[k]);
            g += gq;
          }
          double ieq = id - g * vd;
          four(A, p, m, g);
          if (p >= 0) b.push_back({p, 0, -ieq});
          if (m >= 0) b.push_back({m, 0, ieq});
          break;
        }
        default: stamp_controlled(e, lean, A); break;
      }
    }
    if (!ls.solve(n, A, b, x)) throw RefError{REF_DECOMPOSITION, "clean Newton: LU failed"};
    if (!C.nonlinear) return it + 1;
    bool ok = !any_lim;
    for (int i = 0; i < n && ok; i++) {
      double diff = std::fabs(x[i] - guess[i]);
      double scale = std::max(std::fabs(x[i]), std::fabs(guess[i]));
      double atol = C.names[i][0] == 'I' ? cfg.iabstol : cfg.vabstol;
      ok = diff <= cfg.reltol * scale + atol;
    }
    if (ok) return it + 1;
    guess = x;
  }
  throw RefError{REF_MAX_ITERATIONS, "clean Newton: maximum iterations exceeded"};
}

void trap_update(const Circuit& C, double h, const std::vector<double>& xold, const std::vector<double>& xnew,
                 CleanState& st) {
  for (size_t k = 0; k < C.el.size(); k++) {
    const Elem& e = C.el[k];
    if (e.t == E_C) st.ic[k] = 2.0 * e.value / h * (vdiff(e, xnew) - vdiff(e, xold)) - st.ic[k];
    if (e.t == E_L) st.vl[k] = vdiff(e, xnew);
    if (e.has_jq() && !st.iq.empty())
      st.iq[k] = 2.0 * (junction_charge(e, vdiff(e, xnew)).first - junction_charge(e, vdiff(e, xold)).first) / h -
                 st.iq[k];
  }
}

Circuit* as_c(void* p) { return static_cast<Circuit*>(p); }

template <class F>
int32_t guarded(F&& f) {
  try {
    f();
    return REF_OK;
  } catch (const RefError& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return REF_UNEXPECTED;
  }
}

template <class T>
int dense_lu_solve(int n, std::vector<T>& a, std::vector<T>& b) {
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::abs(a[i * n + k]) > std::abs(a[p * n + k])) p = i;
    if (std::abs(a[p * n + k]) == 0) return REF_DECOMPOSITION;
    if (p != k) {
      for (int j = 0; j < n; j++) std::swap(a[k * n + j], a[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; i++) {
      T l = a[i * n + k] / a[k * n + k];
      if (l == T(0)) continue;
      for (int j = k; j < n; j++) a[i * n + j] -= l * a[k * n + j];
      b[i] -= l * b[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    T s = b[i];
    for (int j = i + 1; j < n; j++) s -= a[i * n + j] * b[j];
    b[i] = s / a[i * n + i];
  }
  return REF_OK;
}

}  // namespace

extern "C" {

void ref_default_config(ref_config* c) {
  c->reltol = 1e-3;
  c->iabstol = 1e-12;
  c->vabstol = 1e-6;
  c->maxiter = 300;
  c->cost = REF_COST_COMPAT;
}

const char* ref_last_error(void) { return g_last_error.c_str(); }

void* ref_parse(const char* netlist, int32_t* status, char* err, int32_t errlen) {
  Circuit* C = nullptr;
  int32_t st = guarded([&] { C = parse_netlist(netlist); });
  if (status) *status = st;
  if (st != REF_OK && err && errlen > 0) {
    std::snprintf(err, errlen, "%s", g_last_error.c_str());
  }
  return C;
}

void* ref_parse_ext(const char* netlist, int32_t flags, int32_t* status, char* err, int32_t errlen) {
  Circuit* C = nullptr;
  int32_t st = guarded([&] { C = parse_netlist(netlist, flags); });
  if (status) *status = st;
  if (st != REF_OK && err && errlen > 0) {
    std::snprintf(err, errlen, "%s", g_last_error.c_str());
  }
  return C;
}

void ref_free(void* c) { delete as_c(c); }
int32_t ref_num_unknowns(void* c) { return N_of(*as_c(c)); }
int32_t ref_num_elements(void* c) { return (int32_t)as_c(c)->el.size(); }
const char* ref_unknown_name(void* c, int32_t i) { return as_c(c)->names.at(i).c_str(); }

int32_t ref_element_info(void* c, int32_t e, char* type, int32_t* plus, int32_t* minus, int32_t* branch,
                         double* value) {
  const Circuit& C = *as_c(c);
  if (e < 0 || e >= (int)C.el.size()) return REF_ELEMENT_NOT_FOUND;
  const Elem& el = C.el[e];
  *type = ET_LETTER[el.t];
  *plus = el.ip;
  *minus = el.im;
  *branch = el.ic;
  *value = el.value;
  return REF_OK;
}

int32_t ref_element_ctrl(void* c, int32_t e, int32_t* ctrl_plus, int32_t* ctrl_minus, int32_t* ctrl_branch) {
  const Circuit& C = *as_c(c);
  if (e < 0 || e >= (int)C.el.size()) return REF_ELEMENT_NOT_FOUND;
  const Elem& el = C.el[e];
  *ctrl_plus = el.icp;
  *ctrl_minus = el.icm;
  *ctrl_branch = el.icb;
  return REF_OK;
}

int32_t ref_set_value(void* c, int32_t e, double v) {
  Circuit& C = *as_c(c);
  if (e < 0 || e >= (int)C.el.size()) return REF_ELEMENT_NOT_FOUND;
  C.el[e].value = v;
  return REF_OK;
}

int32_t ref_op(void* c, const ref_config* cfg, double* x, int32_t* iters) {
  return guarded([&] {
    int it = 0;
    auto v = op_solve(*as_c(c), *cfg, &it);
    std::copy(v.begin(), v.end(), x);
    if (iters) *iters = it;
  });
}

// solver/dc.rs:14-124
int32_t ref_dc(void* c, const ref_config* cfg, const char* element, double start, double stop, double step,
               double* x, int32_t maxpts, int32_t* npts) {
  return guarded([&] {
    Circuit& C = *as_c(c);
    int n = N_of(C);
    int k = -1;
    for (size_t i = 0; i < C.el.size(); i++)
      if (C.el[i].ident() == element) { k = (int)i; break; }
    if (k < 0) throw RefError{REF_ELEMENT_NOT_FOUND, std::string("element not found: ") + element};
    if (C.el[k].t != E_V && C.el[k].t != E_I)
      throw RefError{REF_INVALID_ELEMENT_FORMAT, "DC sweep element must be a voltage or current source"};
    double orig = C.el[k].value;
    size_t num = (size_t)std::fabs((stop - start) / step) + 1;
    if ((int64_t)num > maxpts) throw RefError{REF_UNEXPECTED, "output buffer too small"};
    bool compat = cfg->cost == REF_COST_COMPAT;
    Idx ix{&C, compat};
    LinSolver<double> ls;
    ls.cost = cfg->cost;
    SolMap last(&C, compat);
    std::vector<double> xs;
    for (size_t i = 0; i < num; i++) {
      C.el[k].value = start + (double)i * step;
      SolMap prev = last, res(&C, compat);
      for (int it = 0; it < cfg->maxiter; it++) {
        std::vector<Trip<double>> A, b;
        stamp_dc(C, ix, prev, A, b, true, 0, 0.0, nullptr);
        if (!ls.solve(n, A, b, xs)) {
          C.el[k].value = orig;
          throw RefError{REF_DECOMPOSITION, "DC: LU failed"};
        }
        res.set(xs, "step", (double)i);
        if (!C.nonlinear) break;
        if (converged(prev, res, *cfg)) break;
        prev = res;
        if (it == cfg->maxiter - 1) {
          C.el[k].value = orig;
          throw RefError{REF_MAX_ITERATIONS, "DC: maximum iterations exceeded"};
        }
      }
      last = res;
      std::copy(xs.begin(), xs.end(), x + i * n);
    }
    C.el[k].value = orig;
    *npts = (int32_t)num;
  });
}

// analyses.rs:80-158
int32_t ref_ac_frequencies(int32_t variation, int32_t pts, double fstart, double fstop, double* out,
                           int32_t maxn) {
  std::vector<double> f;
  if (!(fstart <= 0.0 || fstop <= 0.0 || fstart > fstop)) {
    if (variation == 2) {
      if (pts == 1) f.push_back(fstart);
      else if (pts > 1) {
        double step = (fstop - fstart) / (double)(pts - 1);
        for (int i = 0; i < pts; i++) f.push_back(fstart + (double)i * step);
      }
    } else if (pts > 0) {
      double nd = variation == 0 ? std::log10(fstop / fstart) : std::log2(fstop / fstart);
      uint32_t total = (uint32_t)std::round(nd * (double)pts) + 1;
      double factor = std::pow(variation == 0 ? 10.0 : 2.0, 1.0 / (double)pts);
      double cur = fstart;
      for (uint32_t i = 0; i < total; i++) {
        if (cur > fstop * (1.0 + 1e-9)) break;
        f.push_back(cur);
        cur *= factor;
        if (cur > fstop && (f.empty() || f.back() < fstop)) {
          f.push_back(fstop);
          break;
        }
      }
      if (f.empty() || f.back() < fstop * (1.0 - 1e-9)) f.push_back(fstop);
    }
    std::sort(f.begin(), f.end());
    f.erase(std::unique(f.begin(), f.end()), f.end());
  }
  int n = std::min((int)f.size(), (int)maxn);
  std::copy(f.begin(), f.begin() + n, out);
  return (int32_t)f.size();
}

// solver/ac.rs:17-109 (non-positive frequencies are skipped: the caller
// passes only positive ones)
int32_t ref_ac(void* c, const ref_config* cfg, const double* freqs, int32_t nf, double* x) {
  return guarded([&] {
    const Circuit& C = *as_c(c);
    int n = N_of(C);
    auto xop = op_solve(C, *cfg, nullptr);
    bool compat = cfg->cost == REF_COST_COMPAT;
    SolMap op(&C, compat);
    op.set(xop, nullptr, 0.0);
    Idx ix{&C, compat};
    LinSolver<cd> ls;
    ls.cost = cfg->cost;
    std::vector<cd> xs;
    int out = 0;
    for (int k = 0; k < nf; k++) {
      if (freqs[k] <= 0.0) continue;
      std::vector<Trip<cd>> A, b;
      stamp_ac(C, ix, op, freqs[k], A, b);
      if (!ls.solve(n, A, b, xs)) throw RefError{REF_DECOMPOSITION, "AC: LU failed"};
      if (compat) {  // ac.rs:97-103 map build
        std::unordered_map<std::string, cd> m;
        for (auto& kv : C.index) m[kv.first] = xs[kv.second];
        m["frequency"] = cd(freqs[k], 0.0);
      }
      for (int i = 0; i < n; i++) {
        x[(size_t)out * 2 * n + 2 * i] = xs[i].real();
        x[(size_t)out * 2 * n + 2 * i + 1] = xs[i].imag();
      }
      out++;
    }
  });
}

int32_t ref_tran(void* c, const ref_config* cfg, double h, double stop, int32_t method, double* x,
                 int32_t maxpts, int32_t* npts, int32_t* iters) {
  return guarded([&] {
    const Circuit& C = *as_c(c);
    int n = N_of(C);
    size_t num = (size_t)std::round(stop / h);
    if ((int64_t)num + 1 > maxpts) throw RefError{REF_UNEXPECTED, "output buffer too small"};
    *npts = (int32_t)(num + 1);
    if (method == REF_BE_COMPAT) {
      // solver/transient.rs:14-114
      bool compat = cfg->cost == REF_COST_COMPAT;
      Idx ix{&C, compat};
      auto x0 = op_solve(C, *cfg, nullptr);
      SolMap prev(&C, compat);
      prev.set(x0, "time", 0.0);
      std::copy(x0.begin(), x0.end(), x);
      if (iters) iters[0] = 0;
      LinSolver<double> ls;
      ls.cost = cfg->cost;
      std::vector<double> xs;
      for (size_t s = 1; s <= num; s++) {
        double t = (double)s * h;
        SolMap guess = prev, res(&C, compat);
        for (int it = 0; it < cfg->maxiter; it++) {
          std::vector<Trip<double>> A, b;
          stamp_dc(C, ix, guess, A, b, false, 1, h, &prev);
          if (!ls.solve(n, A, b, xs)) throw RefError{REF_DECOMPOSITION, "transient: LU failed"};
          res.set(xs, "time", t);
          if (iters) iters[s] = it + 1;
          if (!C.nonlinear) break;
          if (converged(guess, res, *cfg)) break;
          guess = res;
          if (it == cfg->maxiter - 1) throw RefError{REF_MAX_ITERATIONS, "transient: maximum iterations"};
        }
        prev = res;
        std::copy(xs.begin(), xs.end(), x + s * n);
      }
      return;
    }
    std::vector<double> x0;
    CleanState st;
    st.ic.assign(C.el.size(), 0.0);
    st.vl.assign(C.el.size(), 0.0);
    st.iq.assign(C.el.size(), 0.0);
    clean_newton(C, *cfg, 0, 0.0, 1.0, std::vector<double>(n, 0.0), st, x0);
    // step 0 is the operating point, whose Newton count the OP entry point
    // reports; as for krets' integrator, the step record starts at 0
    if (iters) iters[0] = 0;
    for (size_t k = 0; k < C.el.size(); k++)
      if (C.el[k].t == E_L) st.vl[k] = vdiff(C.el[k], x0);
    std::copy(x0.begin(), x0.end(), x);
    std::vector<double> xp = x0, xn;
    for (size_t s = 1; s <= num; s++) {
      double t = (double)s * h;
      int it = clean_newton(C, *cfg, method == REF_TRAP ? 2 : 1, t, h, xp, st, xn);
      if (iters) iters[s] = it;
      if (method == REF_TRAP) trap_update(C, h, xp, xn, st);
      std::copy(xn.begin(), xn.end(), x + s * n);
      xp = xn;
    }
  });
}

// Backward Euler with divided-difference LTE step control. trace (may be
// null) receives one row per attempted step: t, h, ratio (-1 without an
// estimate), outcome (1 accepted, 0 LTE rejection, -1 Newton failure), next h.
static void tran_lte_run(const Circuit& C, const ref_config& cfg, double tstop, double h0, double hmin, double hmax,
                  double lte_reltol, double lte_abstol, double* times, double* x, int32_t maxpts, int32_t* npts,
                  int32_t* newton_total, int32_t* rejected, std::vector<std::array<double, 5>>* trace) {
  int n = N_of(C);
  CleanState st;
  st.ic.assign(C.el.size(), 0.0);
  st.vl.assign(C.el.size(), 0.0);
  st.iq.assign(C.el.size(), 0.0);
  std::vector<double> x0;
  int nt = clean_newton(C, cfg, 0, 0.0, 1.0, std::vector<double>(n, 0.0), st, x0);
  int rej = 0;
  int cnt = 0;
  auto push = [&](double t, const std::vector<double>& v) {
    if (!times) {
      cnt++;
      return;
    }
    if (cnt >= maxpts) throw RefError{REF_UNEXPECTED, "output buffer too small"};
    times[cnt] = t;
    std::copy(v.begin(), v.end(), x + (size_t)cnt * n);
    cnt++;
  };
  auto note = [&](double t, double h, double ratio, double outcome, double hn) {
    if (trace) trace->push_back({t, h, ratio, outcome, hn});
  };
  push(0.0, x0);
  double t = 0.0, h = h0, tprev = 0.0;
  std::vector<double> xc = x0, xprev = x0, xn;
  bool have_prev = false;
  while (t < tstop * (1.0 - 1e-12)) {
    h = std::min(h, hmax);
    if (t + h > tstop) h = tstop - t;
    int its;
    try {
      its = clean_newton(C, cfg, 1, t + h, h, xc, st, xn);
    } catch (const RefError& e) {
      if (e.code != REF_MAX_ITERATIONS) throw;
      rej++;
      note(t, h, -1.0, -1.0, h / 8.0);
      h /= 8.0;
      if (h < hmin) throw RefError{REF_MAX_ITERATIONS, "LTE: step below hmin after Newton failure"};
      continue;
    }
    nt += its;
    double hnext = h;
    double ratio = -1.0;
    if (have_prev) {
      double hp = t - tprev;
      ratio = 0.0;
      for (const Elem& e : C.el) {
        double a, b0, bm;
        if (e.t == E_C || e.has_jq()) { a = vdiff(e, xn); b0 = vdiff(e, xc); bm = vdiff(e, xprev); }
        else if (e.t == E_L) { a = xn[e.ic]; b0 = xc[e.ic]; bm = xprev[e.ic]; }
        else continue;
        double dd2 = 2.0 * ((a - b0) / h - (b0 - bm) / hp) / (h + hp);
        double lte = 0.5 * h * h * std::fabs(dd2);
        double tol = lte_reltol * std::max(std::fabs(a), std::fabs(b0)) + lte_abstol;
        ratio = std::max(ratio, lte / tol);
      }
      if (ratio > 1.0) {
        rej++;
        double hr = h * std::max(0.25, 0.9 / std::sqrt(ratio));
        note(t, h, ratio, 0.0, hr);
        h = hr;
        if (h < hmin) throw RefError{REF_MAX_ITERATIONS, "LTE: step below hmin"};
        continue;
      }
      hnext = ratio > 0 ? h * std::min(2.0, std::max(0.25, 0.9 / std::sqrt(ratio))) : 2.0 * h;
    }
    note(t, h, ratio, 1.0, hnext);
    tprev = t;
    xprev = xc;
    t += h;
    xc = xn;
    have_prev = true;
    push(t, xc);
    h = hnext;
  }
  *npts = cnt;
  if (newton_total) *newton_total = nt;
  if (rejected) *rejected = rej;
}

int32_t ref_tran_lte(void* c, const ref_config* cfg, double tstop, double h0, double hmin, double hmax,
                     double lte_reltol, double lte_abstol, double* times, double* x, int32_t maxpts,
                     int32_t* npts, int32_t* newton_total, int32_t* rejected) {
  return guarded([&] {
    tran_lte_run(*as_c(c), *cfg, tstop, h0, hmin, hmax, lte_reltol, lte_abstol, times, x, maxpts, npts,
                 newton_total, rejected, nullptr);
  });
}

int32_t ref_tran_lte_trace(void* c, const ref_config* cfg, double tstop, double h0, double hmin, double hmax,
                           double lte_reltol, double lte_abstol, double* trace, int32_t maxtr, int32_t* ntr) {
  return guarded([&] {
    std::vector<std::array<double, 5>> tr;
    int32_t npts = 0;
    tran_lte_run(*as_c(c), *cfg, tstop, h0, hmin, hmax, lte_reltol, lte_abstol, nullptr, nullptr, 0, &npts,
                 nullptr, nullptr, &tr);
    if ((int64_t)tr.size() > maxtr) throw RefError{REF_UNEXPECTED, "trace buffer too small"};
    for (size_t k = 0; k < tr.size(); k++) std::copy(tr[k].begin(), tr[k].end(), trace + 5 * k);
    *ntr = (int32_t)tr.size();
  });
}

// Clean backward Euler on a given time grid: the step of ref_tran_lte
// (clean_newton with the source at the step's end) without its controller.
// times[0] must be 0; x[0] is the clean operating point.
int32_t ref_tran_grid(void* c, const ref_config* cfg, const double* times, int32_t npts, double* x,
                      int32_t* newton_total) {
  return guarded([&] {
    const Circuit& C = *as_c(c);
    const int n = N_of(C);
    if (npts < 1 || times[0] != 0.0) throw RefError{REF_UNEXPECTED, "the grid must start at t = 0"};
    CleanState st;
    st.ic.assign(C.el.size(), 0.0);
    st.vl.assign(C.el.size(), 0.0);
    st.iq.assign(C.el.size(), 0.0);
    std::vector<double> xs;
    int nt = clean_newton(C, *cfg, 0, 0.0, 1.0, std::vector<double>(n, 0.0), st, xs);
    std::copy(xs.begin(), xs.end(), x);
    for (int32_t s = 1; s < npts; s++) {
      const double h = times[s] - times[s - 1];
      if (!(h > 0.0)) throw RefError{REF_UNEXPECTED, "the grid must increase"};
      std::vector<double> xp(x + (size_t)(s - 1) * n, x + (size_t)s * n), xn;
      nt += clean_newton(C, *cfg, 1, times[s], h, xp, st, xn);
      std::copy(xn.begin(), xn.end(), x + (size_t)s * n);
    }
    if (newton_total) *newton_total = nt;
  });
}

double ref_wave_value(int32_t kind, const double* p, double t) {
  Elem e{};
  e.t = E_V;
  e.w.kind = kind;
  std::copy(p, p + 7, e.w.p);
  e.value = p[0];
  return wave_value(e, t);
}

int32_t ref_parse_value(const char* tok, double* v) {
  double out = 0.0;
  if (!tok || !parse_value(tok, out)) return REF_PARSE;
  *v = out;
  return REF_OK;
}

int32_t ref_pnjlim(double vnew, double vold, double vt, double vcrit, double* out) {
  bool lim = false;
  *out = pnjlim(vnew, vold, vt, vcrit, lim);
  return lim ? 1 : 0;
}

int32_t ref_tran_adjoint(void* c, const ref_config* cfg, double h, int32_t N, int32_t method,
                         const double* cvec, double* J, double* grad) {
  return guarded([&] {
    const Circuit& C = *as_c(c);
    if (C.nonlinear) throw RefError{REF_UNEXPECTED, "adjoint: linear circuits only"};
    int n = N_of(C);
    int ne = (int)C.el.size();
    double alpha = method == REF_TRAP ? 2.0 : 1.0;
    bool trap = method == REF_TRAP;
    // forward pass, all states kept (test sizes only)
    std::vector<std::vector<double>> X(N + 1);
    CleanState st;
    st.ic.assign(ne, 0.0);
    st.vl.assign(ne, 0.0);
    clean_newton(C, *cfg, 0, 0.0, 1.0, std::vector<double>(n, 0.0), st, X[0]);
    for (int k = 0; k < ne; k++)
      if (C.el[k].t == E_L) st.vl[k] = vdiff(C.el[k], X[0]);
    for (int s = 1; s <= N; s++) {
      clean_newton(C, *cfg, trap ? 2 : 1, s * h, h, X[s - 1], st, X[s]);
      if (trap) trap_update(C, h, X[s - 1], X[s], st);
    }
    auto wt = [&](int s) { return (s == 0 || s == N) ? 0.5 * h : h; };
    double Jv = 0.0;
    for (int s = 0; s <= N; s++)
      for (int i = 0; i < n; i++) Jv += wt(s) * cvec[i] * X[s][i];
    *J = Jv;
    // constant step matrix A and its LU
    std::vector<Trip<double>> A, Adc;
    for (const Elem& e : C.el) {
      int p = e.ip, m = e.im, cc = e.ic;
      if (e.t == E_R) { four(A, p, m, 1.0 / e.value); four(Adc, p, m, 1.0 / e.value); }
      if (e.t == E_C) four(A, p, m, alpha * e.value / h);
      if (e.t == E_L || e.t == E_V) {
        for (auto* T : {&A, &Adc}) {
          if (p >= 0) { T->push_back({p, cc, 1.0}); T->push_back({cc, p, 1.0}); }
          if (m >= 0) { T->push_back({m, cc, -1.0}); T->push_back({cc, m, -1.0}); }
        }
        if (e.t == E_L) A.push_back({cc, cc, -alpha * e.value / h});
      }
      if (e.t == E_I) {
        for (auto* T : {&A, &Adc}) {
          T->push_back({cc, cc, 1.0});
          if (p >= 0) T->push_back({p, cc, 1.0});
          if (m >= 0) T->push_back({m, cc, -1.0});
        }
      }
      if (e.t >= E_E) {  // controlled sources: constant; their gains are not differentiated
        stamp_controlled(e, Idx{&C, false}, A);
        stamp_controlled(e, Idx{&C, false}, Adc);
      }
    }
    auto factor = [&](std::vector<Trip<double>>& T, SpLU<double>& lu) {
      std::vector<int> cp, ri;
      std::vector<double> vx;
      sum_to_csc(n, T, REF_COST_LEAN, cp, ri, vx);
      if (!lu.factor(n, cp, ri, vx, min_degree(n, cp, ri))) throw RefError{REF_DECOMPOSITION, "adjoint LU"};
    };
    SpLU<double> lu, ludc;
    factor(A, lu);
    factor(Adc, ludc);
    std::vector<double> g(ne, 0.0);
    std::vector<double> xbar(n), icbar(ne, 0.0), vlbar(ne, 0.0), lam;
    for (int i = 0; i < n; i++) xbar[i] = wt(N) * cvec[i];
    auto ev = [&](const Elem& e, const std::vector<double>& v) { return vdiff(e, v); };
    for (int s = N - 1; s >= 0; s--) {
      const auto& x1 = X[s + 1];
      const auto& x0 = X[s];
      std::vector<double> xb0(n, 0.0), icb0(ne, 0.0), vlb0(ne, 0.0);
      if (trap) {
        for (int k = 0; k < ne; k++) {
          const Elem& e = C.el[k];
          if (e.t == E_C && icbar[k] != 0.0) {
            double a = 2.0 * e.value / h * icbar[k];
            if (e.ip >= 0) { xbar[e.ip] += a; xb0[e.ip] -= a; }
            if (e.im >= 0) { xbar[e.im] -= a; xb0[e.im] += a; }
            icb0[k] -= icbar[k];
            g[k] += icbar[k] * 2.0 / h * (ev(e, x1) - ev(e, x0));
          }
          if (e.t == E_L && vlbar[k] != 0.0) {
            if (e.ip >= 0) xbar[e.ip] += vlbar[k];
            if (e.im >= 0) xbar[e.im] -= vlbar[k];
          }
        }
      }
      lu.solve_t(xbar, lam);
      for (int k = 0; k < ne; k++) {
        const Elem& e = C.el[k];
        if (e.t == E_R) {
          double el = ev(e, lam), ex = ev(e, x1);
          g[k] += el * ex / (e.value * e.value);
        } else if (e.t == E_C) {
          double el = ev(e, lam);
          double a = alpha * e.value / h * el;
          if (e.ip >= 0) xb0[e.ip] += a;
          if (e.im >= 0) xb0[e.im] -= a;
          if (trap) icb0[k] += el;
          g[k] += el * alpha / h * ev(e, x0) - alpha / h * el * ev(e, x1);
        } else if (e.t == E_L) {
          double lb = lam[e.ic];
          xb0[e.ic] += -alpha * e.value / h * lb;
          if (trap) vlb0[k] += -lb;
          g[k] += lb * (-alpha / h) * x0[e.ic] + alpha / h * lb * x1[e.ic];
        }
      }
      for (int i = 0; i < n; i++) xb0[i] += wt(s) * cvec[i];
      xbar = xb0;
      icbar = icb0;
      vlbar = vlb0;
    }
    // x_0 = OP: vl_0 depends on x_0; ic_0 is a constant zero.
    for (int k = 0; k < ne; k++) {
      const Elem& e = C.el[k];
      if (e.t == E_L && trap && vlbar[k] != 0.0) {
        if (e.ip >= 0) xbar[e.ip] += vlbar[k];
        if (e.im >= 0) xbar[e.im] -= vlbar[k];
      }
    }
    ludc.solve_t(xbar, lam);
    for (int k = 0; k < ne; k++) {
      const Elem& e = C.el[k];
      if (e.t == E_R) g[k] += ev(e, lam) * ev(e, X[0]) / (e.value * e.value);
    }
    std::copy(g.begin(), g.end(), grad);
  });
}

int32_t ref_dense_solve(int32_t n, const double* A, const double* b, double* x) {
  std::vector<long double> a(A, A + (size_t)n * n), bb(b, b + n);
  int r = dense_lu_solve<long double>(n, a, bb);
  for (int i = 0; i < n; i++) x[i] = (double)bb[i];
  return r;
}

int32_t ref_dense_solve_c(int32_t n, const double* A, const double* b, double* x) {
  using cl = std::complex<long double>;
  std::vector<cl> a((size_t)n * n), bb(n);
  for (size_t i = 0; i < (size_t)n * n; i++) a[i] = cl(A[2 * i], A[2 * i + 1]);
  for (int i = 0; i < n; i++) bb[i] = cl(b[2 * i], b[2 * i + 1]);
  int r = dense_lu_solve<cl>(n, a, bb);
  for (int i = 0; i < n; i++) {
    x[2 * i] = (double)bb[i].real();
    x[2 * i + 1] = (double)bb[i].imag();
  }
  return r;
}

int32_t ref_assemble_dc(void* c, const double* guess, double* A, double* b) {
  return guarded([&] {
    const Circuit& C = *as_c(c);
    int n = N_of(C);
    SolMap s(&C, false);
    if (guess) s.set(std::vector<double>(guess, guess + n), nullptr, 0.0);
    std::vector<Trip<double>> T, B;
    stamp_dc(C, Idx{&C, false}, s, T, B, true, 0, 0.0, nullptr);
    std::fill(A, A + (size_t)n * n, 0.0);
    std::fill(b, b + n, 0.0);
    for (auto& t : T) A[(size_t)t.r * n + t.c] += t.v;
    for (auto& t : B) b[t.r] += t.v;
  });
}

int32_t ref_assemble_ac(void* c, const double* xop, double f, double* A, double* b) {
  return guarded([&] {
    const Circuit& C = *as_c(c);
    int n = N_of(C);
    SolMap s(&C, false);
    s.set(std::vector<double>(xop, xop + n), nullptr, 0.0);
    std::vector<Trip<cd>> T, B;
    stamp_ac(C, Idx{&C, false}, s, f, T, B);
    std::fill(A, A + (size_t)2 * n * n, 0.0);
    std::fill(b, b + 2 * n, 0.0);
    for (auto& t : T) {
      A[2 * ((size_t)t.r * n + t.c)] += t.v.real();
      A[2 * ((size_t)t.r * n + t.c) + 1] += t.v.imag();
    }
    for (auto& t : B) {
      b[2 * t.r] += t.v.real();
      b[2 * t.r + 1] += t.v.imag();
    }
  });
}

}  // extern "C"
