// Netlist reader for the R/L/C/V/I/D subset of krets' grammar, producing the
// element tables of the C ABI in krets' unknown numbering.
//
// Follows crates/krets-parser/src/parser.rs:27-154 (line handling, .control
// skipping, subcircuit expansion, .model binding, index map at 105-125), the
// element parsers in elements/{resistor,capacitor,inductor,current_source,
// diode,voltage_source}.rs, utils.rs:32-78 (SI suffixes, value tokens),
// models.rs:40-95 and elements/subcircuit.rs:33-200 (definitions collected in
// a first pass, instances expanded in place: pins map to the instance's
// nodes, every other node, ground "0" included, and every element name get
// the instance name and "_" as prefix, nested instances recursively).
// Elements are matched in krets' `alt` order (elements.rs:44-65) and, like
// krets, a line is not required to be consumed entirely.
//
// Beyond krets' grammar: the linear controlled sources of its README to-do
// list (README.md:56-62), in SPICE syntax. "E<name> n+ n- nc+ nc- gain"
// (VCVS) and "H<name> n+ n- <source> r" (CCVS) own a branch current I(<id>)
// like a voltage source; "G<name> n+ n- nc+ nc- gm" (VCCS) and
// "F<name> n+ n- <source> gain" (CCCS) do not. Values take SI suffixes. In
// the index map the controlling nodes of E and G follow n+ and n-. <source>
// names an element with a branch current (V, L, I, E, H or a G2 resistor);
// inside a subcircuit it names that instance's own element.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "krets_gpu.h"

namespace {

struct Rec {
  char letter;
  std::string name, plus, minus, model;
  std::string cplus, cminus, ctrl;  // E/G controlling nodes, F/H controlling element
  double p[KG_NPARAM] = {0};
  bool g2 = false;
};

struct Scan {
  const std::string* s;
  size_t i = 0;
  bool end() const { return i >= s->size(); }
  char at() const { return end() ? '\0' : (*s)[i]; }
};

bool blank(char c) { return c == ' ' || c == '\t'; }

bool sp1(Scan& c) {
  size_t j = c.i;
  while (j < c.s->size() && blank((*c.s)[j])) j++;
  if (j == c.i) return false;
  c.i = j;
  return true;
}
void sp0(Scan& c) {
  while (!c.end() && blank(c.at())) c.i++;
}
bool word(Scan& c, const char* w, bool nocase) {
  size_t n = std::strlen(w);
  if (c.i + n > c.s->size()) return false;
  for (size_t k = 0; k < n; k++) {
    char a = (*c.s)[c.i + k], b = w[k];
    if (nocase ? std::tolower((unsigned char)a) != std::tolower((unsigned char)b) : a != b) return false;
  }
  c.i += n;
  return true;
}
bool ident(Scan& c, std::string& out) {
  size_t j = c.i;
  while (j < c.s->size() && (std::isalnum((unsigned char)(*c.s)[j]) || (*c.s)[j] == '_')) j++;
  if (j == c.i) return false;
  out = c.s->substr(c.i, j - c.i);
  c.i = j;
  return true;
}

// Rust f64::from_str: [+-](inf|infinity|nan|digits[.digits][e[+-]digits]|.digits[...])
bool f64_text(const std::string& t, double& v) {
  std::string u;
  for (char ch : t) u.push_back((char)std::tolower((unsigned char)ch));
  size_t k = (!u.empty() && (u[0] == '+' || u[0] == '-')) ? 1 : 0;
  std::string r = u.substr(k);
  if (r == "inf" || r == "infinity" || r == "nan") {
    v = std::strtod(u.c_str(), nullptr);
    return true;
  }
  size_t a = k, ndig = 0;
  while (a < u.size() && std::isdigit((unsigned char)u[a])) a++, ndig++;
  if (a < u.size() && u[a] == '.') {
    a++;
    while (a < u.size() && std::isdigit((unsigned char)u[a])) a++, ndig++;
  }
  if (!ndig) return false;
  if (a < u.size() && u[a] == 'e') {
    a++;
    if (a < u.size() && (u[a] == '+' || u[a] == '-')) a++;
    size_t e0 = a;
    while (a < u.size() && std::isdigit((unsigned char)u[a])) a++;
    if (a == e0) return false;
  }
  if (a != u.size()) return false;
  v = std::strtod(u.c_str(), nullptr);
  return true;
}

// utils.rs:32-64 (parse_value) applied to a value token (utils.rs:72-78)
bool si_value(Scan& c, double& v) {
  size_t j = c.i;
  while (j < c.s->size() && !std::strchr(" \t\r\n()", (*c.s)[j])) j++;
  if (j == c.i) return false;
  std::string tok;
  for (size_t k = c.i; k < j; k++) tok.push_back((char)std::toupper((unsigned char)(*c.s)[k]));
  double mul = 1.0;
  if (tok.size() >= 3 && tok.compare(tok.size() - 3, 3, "MEG") == 0) {
    mul = 1e6;
    tok.resize(tok.size() - 3);
  } else if (!tok.empty()) {
    static const char* suf = "FPNUMKGT";
    static const double mult[] = {1e-15, 1e-12, 1e-9, 1e-6, 1e-3, 1e3, 1e9, 1e12};
    const char* q = std::strchr(suf, tok.back());
    if (q && *q) {
      mul = mult[q - suf];
      tok.pop_back();
    }
  }
  double base;
  if (!f64_text(tok, base)) return false;
  v = base * mul;
  c.i = j;
  return true;
}

// nom::number::complete::double (used for V source dc/ac values).
// 1 = parsed, 0 = no match, -1 = exponent without digits (hard error).
int nom_float(Scan& c, double& v) {
  const std::string& s = *c.s;
  size_t j = c.i;
  if (j < s.size() && (s[j] == '+' || s[j] == '-')) j++;
  size_t d = j;
  bool num = false;
  while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
  if (j > d) {
    num = true;
    if (j < s.size() && s[j] == '.') {
      j++;
      while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
    }
  } else if (j < s.size() && s[j] == '.' && j + 1 < s.size() && std::isdigit((unsigned char)s[j + 1])) {
    num = true;
    j++;
    while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
  }
  if (num) {
    if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
      size_t q = j + 1;
      if (q < s.size() && (s[q] == '+' || s[q] == '-')) q++;
      size_t e0 = q;
      while (q < s.size() && std::isdigit((unsigned char)s[q])) q++;
      if (q == e0) return -1;
      j = q;
    }
    v = std::strtod(s.substr(c.i, j - c.i).c_str(), nullptr);
    c.i = j;
    return 1;
  }
  for (const char* w : {"nan", "infinity", "inf"}) {
    Scan t = c;
    if (word(t, w, true)) {
      v = std::strtod(w, nullptr);
      c.i = t.i;
      return 1;
    }
  }
  return 0;
}

bool head(Scan& c, char letter, Rec& r) {
  char w[2] = {letter, 0};
  r = Rec();
  r.letter = letter;
  return word(c, w, true) && ident(c, r.name) && sp1(c) && ident(c, r.plus) && sp1(c) && ident(c, r.minus);
}

bool parse_two_terminal(const std::string& line, char letter, Rec& r) {
  Scan c{&line};
  if (!head(c, letter, r) || !sp1(c) || !si_value(c, r.p[0])) return false;
  if (letter == 'C') {
    Scan d = c;
    if (sp1(d) && word(d, "G2", true)) r.g2 = true;
  }
  return true;
}

// E/G: "n+ n- nc+ nc- value"; F/H: "n+ n- <source> value"
bool parse_controlled(const std::string& line, char letter, Rec& r) {
  Scan c{&line};
  if (!head(c, letter, r) || !sp1(c)) return false;
  if (letter == 'E' || letter == 'G') {
    if (!ident(c, r.cplus) || !sp1(c) || !ident(c, r.cminus) || !sp1(c)) return false;
  } else if (!ident(c, r.ctrl) || !sp1(c)) {
    return false;
  }
  return si_value(c, r.p[0]);
}

bool parse_diode(const std::string& line, Rec& r) {
  Scan c{&line};
  if (!head(c, 'D', r)) return false;
  Scan d = c;
  std::string m;
  r.model = (sp1(d) && ident(d, m)) ? m : "default";
  return true;
}

bool parse_wave(Scan& c, const char* tag, int nv, double* out) {
  Scan d = c;
  if (!word(d, tag, true)) return false;
  sp0(d);
  if (!word(d, "(", false)) return false;
  for (int k = 0; k < nv; k++) {
    if (k == 0) sp0(d);
    else if (!sp1(d)) return false;
    if (!si_value(d, out[k])) return false;
  }
  sp0(d);
  if (!word(d, ")", false)) return false;
  c = d;
  return true;
}

// voltage_source.rs:187-235: optional implicit dc, then dc/ac/pulse/sin blocks
int parse_vsrc(const std::string& line, Rec& r) {
  Scan c{&line};
  if (!head(c, 'V', r)) return 0;
  {
    Scan d = c;
    double v;
    if (sp1(d)) {
      int k = nom_float(d, v);
      if (k < 0) return -1;
      if (k > 0) {
        r.p[0] = v;
        c = d;
      }
    }
  }
  for (;;) {
    Scan d = c;
    if (!sp1(d)) break;
    double v, w[7];
    Scan a = d;
    if (word(a, "dc", true) && sp1(a)) {
      int k = nom_float(a, v);
      if (k < 0) return -1;
      if (k > 0) { r.p[0] = v; c = a; continue; }
    }
    a = d;
    if (word(a, "ac", true) && sp1(a)) {
      int k = nom_float(a, v);
      if (k < 0) return -1;
      if (k > 0) { r.p[1] = v; c = a; continue; }
    }
    a = d;
    if (parse_wave(a, "pulse", 7, w)) {
      r.p[2] = KG_WAVE_PULSE;
      for (int k = 0; k < 7; k++) r.p[3 + k] = w[k];
      c = a;
      continue;
    }
    a = d;
    if (parse_wave(a, "sin", 6, w)) {
      r.p[2] = KG_WAVE_SIN;
      for (int k = 0; k < 7; k++) r.p[3 + k] = k < 6 ? w[k] : 0.0;
      c = a;
      continue;
    }
    break;
  }
  return 1;
}

// Diode model (models/diode.rs:14-37: is, rs, n; krets ignores other keys).
// cjo (cj0), vj, m, fc and tt are the SPICE junction-charge parameters,
// read for the opt-in diode parasitics of kg_parse_netlist_ext only.
struct Model {
  bool diode;
  double is = 1e-12, n = 1.0, rs = 0.0;
  double cjo = 0.0, vj = 1.0, m = 0.5, fc = 0.5, tt = 0.0;
};

bool parse_model(std::string line, std::string& name, Model& m) {
  line = line.substr(0, line.find('%'));
  size_t a = line.find_first_not_of(" \t\r\n"), b = line.find_last_not_of(" \t\r\n");
  line = a == std::string::npos ? std::string() : line.substr(a, b - a + 1);
  Scan c{&line};
  if (!word(c, ".model", true) || !sp1(c) || !ident(c, name) || !sp1(c)) return false;
  if (word(c, "NMOS", false) || word(c, "PMOS", false)) m.diode = false;
  else if (word(c, "D", false)) m.diode = true;
  else return false;
  sp0(c);
  word(c, "(", false);
  for (;;) {
    Scan d = c;
    std::string k;
    double v;
    sp0(d);
    if (!ident(d, k)) break;
    sp0(d);
    if (!word(d, "=", false)) break;
    sp0(d);
    if (!si_value(d, v)) break;
    for (auto& ch : k) ch = (char)std::tolower((unsigned char)ch);
    if (k == "is") m.is = v;
    else if (k == "n") m.n = v;
    else if (k == "rs") m.rs = v;
    else if (k == "cjo" || k == "cj0") m.cjo = v;
    else if (k == "vj") m.vj = v;
    else if (k == "m") m.m = v;
    else if (k == "fc") m.fc = v;
    else if (k == "tt") m.tt = v;
    c = d;
  }
  return true;
}

std::string low(const std::string& s) {
  std::string r(s);
  for (auto& ch : r) ch = (char)std::tolower((unsigned char)ch);
  return r;
}

// ---- subcircuits (elements/subcircuit.rs)
struct XInst {
  std::string iname, def;
  std::vector<std::string> nodes;
};
struct SubItem {
  bool is_x = false;
  Rec r;
  XInst x;
};
struct SubDef {
  std::vector<std::string> pins;
  std::vector<SubItem> items;
};

// space1, then an identifier ([A-Za-z0-9_]+); false (c untouched) if absent
bool next_ident(Scan& c, std::string& out) {
  Scan d = c;
  if (!sp1(d) || !ident(d, out)) return false;
  c = d;
  return true;
}

// ".subckt NAME pin..." (subcircuit.rs:128-132)
bool parse_subckt_header(const std::string& line, std::string& name, std::vector<std::string>& pins) {
  Scan c{&line};
  if (!word(c, ".subckt", true) || !next_ident(c, name)) return false;
  std::string p;
  while (next_ident(c, p)) pins.push_back(p);
  return true;
}

// "X<name> node... DEF" (subcircuit.rs:134-157)
bool parse_xinst(const std::string& line, XInst& x) {
  Scan c{&line};
  if (!word(c, "x", true) || !ident(c, x.iname)) return false;
  std::vector<std::string> toks;
  std::string t;
  while (next_ident(c, t)) toks.push_back(t);
  if (toks.empty()) return false;
  x.def = toks.back();
  toks.pop_back();
  x.nodes = toks;
  return true;
}

}  // namespace

struct kg_circuit_impl {
  std::vector<Rec> recs;
  std::vector<kg_element> el;
  std::vector<std::string> names;
  std::vector<std::string> ids;
};

extern "C" int32_t kg_parse_netlist(const char* text, kg_circuit** out, char* err, int32_t errlen) {
  return kg_parse_netlist_ext(text, 0, out, err, errlen);
}

extern "C" int32_t kg_parse_netlist_ext(const char* text, int32_t flags, kg_circuit** out, char* err,
                                        int32_t errlen) {
  auto fail = [&](int32_t code, const std::string& msg) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", msg.c_str());
    return code;
  };
  if (!text || !out) return fail(KG_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  std::string src(text);
  auto trimmed = [](std::string line) {
    size_t a = line.find_first_not_of(" \t\r\n\v\f"), b = line.find_last_not_of(" \t\r\n\v\f");
    return a == std::string::npos ? std::string() : line.substr(a, b - a + 1);
  };
  // one primitive element line, krets' alt order; 0 no match, 1 parsed,
  // -1 a V line that is not a valid source
  auto primitive = [](const std::string& line, Rec& r) {
    int vs = 0;
    bool ok = parse_two_terminal(line, 'R', r) || parse_two_terminal(line, 'C', r) ||
              parse_two_terminal(line, 'L', r) || (vs = parse_vsrc(line, r)) == 1;
    if (!ok && vs == 0) ok = parse_two_terminal(line, 'I', r) || parse_diode(line, r);
    if (!ok && vs == 0)
      for (char k : {'E', 'G', 'F', 'H'})
        if (!ok) ok = parse_controlled(line, k, r);
    return ok ? 1 : 0;
  };
  // first pass (subcircuit.rs:159-200): definitions anywhere in the text
  std::map<std::string, SubDef> defs;
  {
    std::string cur;
    bool in_def = false;
    size_t p0 = 0;
    while (p0 < src.size()) {
      size_t e = src.find('\n', p0);
      if (e == std::string::npos) e = src.size();
      std::string line = trimmed(src.substr(p0, e - p0));
      p0 = e + 1;
      if (line.empty() || line[0] == '%' || line[0] == '*') continue;
      std::string lo = low(line);
      if (lo.rfind(".subckt", 0) == 0) {
        std::string nm;
        std::vector<std::string> pins;
        if (!parse_subckt_header(line, nm, pins)) return fail(KG_PARSE, "failed to parse subcircuit header: " + line);
        defs[nm] = SubDef{pins, {}};
        cur = nm;
        in_def = true;
        continue;
      }
      if (lo.rfind(".ends", 0) == 0) {
        in_def = false;
        cur.clear();
        continue;
      }
      if (!in_def) continue;
      SubItem it;
      if (primitive(line, it.r) != 1) {
        if (!parse_xinst(line, it.x)) {
          char f = (char)std::toupper((unsigned char)line[0]);
          return fail(f == 'Q' || f == 'M' ? KG_INVALID_ELEMENT_FORMAT : KG_PARSE,
                      "failed to parse subcircuit element in '" + cur + "': " + line);
        }
        it.is_x = true;
      }
      defs[cur].items.push_back(it);
    }
  }
  // instance expansion (subcircuit.rs:33-110), in place
  std::function<bool(const XInst&, std::vector<Rec>&, int, std::string&)> expand =
      [&](const XInst& x, std::vector<Rec>& out, int depth, std::string& why) -> bool {
    auto d = defs.find(x.def);
    if (d == defs.end()) {
      why = "undefined subcircuit definition: " + x.def;
      return false;
    }
    if (x.nodes.size() != d->second.pins.size()) {
      why = "node/port mismatch for instance " + x.iname + ": expected " + std::to_string(d->second.pins.size()) +
            ", got " + std::to_string(x.nodes.size());
      return false;
    }
    if (depth > 64) {
      why = "subcircuit nesting deeper than 64 levels (recursive definition?): " + x.def;
      return false;
    }
    std::map<std::string, std::string> p2n;  // a repeated pin maps to its last node
    for (size_t k = 0; k < x.nodes.size(); k++) p2n[d->second.pins[k]] = x.nodes[k];
    auto map = [&](const std::string& nd) {
      auto f = p2n.find(nd);
      return f != p2n.end() ? f->second : x.iname + "_" + nd;
    };
    for (const SubItem& it : d->second.items) {
      if (!it.is_x) {
        Rec r = it.r;
        r.plus = map(r.plus);
        r.minus = map(r.minus);
        if (r.letter == 'E' || r.letter == 'G') {
          r.cplus = map(r.cplus);
          r.cminus = map(r.cminus);
        }
        if (!r.ctrl.empty()) r.ctrl = r.ctrl.substr(0, 1) + x.iname + "_" + r.ctrl.substr(1);
        r.name = x.iname + "_" + r.name;
        out.push_back(r);
      } else {
        XInst y = it.x;
        for (auto& nd : y.nodes) nd = map(nd);
        y.iname = x.iname + "_" + y.iname;
        if (!expand(y, out, depth + 1, why)) return false;
      }
    }
    return true;
  };
  auto* C = new kg_circuit_impl();
  std::map<std::string, Model> models;
  bool ctl = false, sub = false;
  size_t pos = 0;
  int lineno = 0;
  while (pos < src.size()) {
    size_t e = src.find('\n', pos);
    if (e == std::string::npos) e = src.size();
    std::string line = src.substr(pos, e - pos);
    pos = e + 1;
    lineno++;
    size_t a = line.find_first_not_of(" \t\r\n\v\f"), b = line.find_last_not_of(" \t\r\n\v\f");
    line = a == std::string::npos ? std::string() : line.substr(a, b - a + 1);
    if (line.empty() || line[0] == '%' || line[0] == '*') continue;
    std::string lo = low(line);
    if (lo.rfind(".control", 0) == 0) { ctl = true; continue; }
    if (lo.rfind(".endc", 0) == 0) { ctl = false; continue; }
    if (ctl) continue;
    if (lo.rfind(".subckt", 0) == 0) { sub = true; continue; }
    if (lo.rfind(".ends", 0) == 0) { sub = false; continue; }
    if (sub) continue;
    if (lo == ".end") continue;
    if (lo.rfind(".model", 0) == 0) {
      std::string nm;
      Model m;
      if (!parse_model(line, nm, m)) {
        delete C;
        return fail(KG_PARSE, "line " + std::to_string(lineno) + ": invalid .model card");
      }
      models[nm] = m;
      continue;
    }
    Rec r;
    if (primitive(line, r) != 1) {
      XInst x;
      if (parse_xinst(line, x)) {
        std::string why;
        if (!expand(x, C->recs, 0, why)) {
          delete C;
          return fail(KG_PARSE, "line " + std::to_string(lineno) + ": " + why);
        }
        continue;
      }
      char f = (char)std::toupper((unsigned char)line[0]);
      delete C;
      if (f == 'Q' || f == 'M')
        return fail(KG_INVALID_ELEMENT_FORMAT,
                    "line " + std::to_string(lineno) + ": BJT/MOSFET elements are not supported");
      return fail(KG_PARSE, "line " + std::to_string(lineno) + ": cannot parse '" + line + "'");
    }
    C->recs.push_back(r);
  }
  if (C->recs.empty()) {
    delete C;
    return fail(KG_EMPTY_NETLIST, "empty netlist");
  }
  // parser.rs:105-125: branch current first (group 2), then new nodes in
  // (plus, minus) order; ground "0" never gets an index.
  std::map<std::string, int> idx;
  std::set<std::string> seen;
  int counter = 0;
  auto g2 = [](const Rec& r) {
    return r.letter == 'V' || r.letter == 'L' || r.letter == 'I' || r.letter == 'E' || r.letter == 'H' ||
           ((r.letter == 'C' || r.letter == 'R') && r.g2);
  };
  for (auto& r : C->recs) {
    std::string id = std::string(1, r.letter) + r.name;
    C->ids.push_back(id);
    if (g2(r)) {
      std::string k = "I(" + id + ")";
      if (idx.count(k)) {
        delete C;
        return fail(KG_UNEXPECTED, "duplicate element identifier " + id);
      }
      idx[k] = counter++;
    }
    const bool vc = r.letter == 'E' || r.letter == 'G';
    for (const std::string* nd : {&r.plus, &r.minus, &r.cplus, &r.cminus}) {
      if (!vc && (nd == &r.cplus || nd == &r.cminus)) continue;
      if (seen.insert(*nd).second && *nd != "0") idx["V(" + *nd + ")"] = counter++;
    }
  }
  C->names.assign(counter, "");
  for (auto& kv : idx) C->names[kv.second] = kv.first;
  for (size_t k = 0; k < C->recs.size(); k++) {
    Rec& r = C->recs[k];
    if (r.letter == 'D') {  // parser.rs:131-140: a diode needs its .model
      auto it = models.find(r.model);
      if (it == models.end() || !it->second.diode) {
        std::string msg = "diode " + C->ids[k] + " refers to missing .model '" + r.model + "'";
        delete C;
        return fail(KG_UNEXPECTED, msg);
      }
      r.p[0] = it->second.is;
      r.p[1] = it->second.n;
      if (flags & KG_NETLIST_DIODE_PARASITICS) {
        const Model& md = it->second;
        r.p[2] = md.cjo;
        r.p[3] = md.vj;
        r.p[4] = md.m;
        r.p[5] = md.fc;
        r.p[6] = md.tt;
        r.p[7] = md.rs;  // consumed below: the series resistor element
      }
    }
    kg_element e{};
    static const char* L = "RCLVIDEGFH";
    e.type = (int32_t)(std::strchr(L, r.letter) - L);
    e.ctrl_plus = e.ctrl_minus = e.ctrl_branch = -1;
    auto node = [&](const std::string& nd) {
      auto it = idx.find("V(" + nd + ")");
      return it == idx.end() ? -1 : it->second;
    };
    e.plus = node(r.plus);
    e.minus = node(r.minus);
    auto it = idx.find("I(" + C->ids[k] + ")");
    e.branch = it == idx.end() ? -1 : it->second;
    e.g2 = r.g2 ? 1 : 0;
    if (r.letter == 'E' || r.letter == 'G') {
      e.ctrl_plus = node(r.cplus);
      e.ctrl_minus = node(r.cminus);
    } else if (r.letter == 'F' || r.letter == 'H') {
      std::string id = r.ctrl;
      id[0] = (char)std::toupper((unsigned char)id[0]);
      auto b = idx.find("I(" + id + ")");
      if (b == idx.end()) {
        std::string msg = "controlled source " + C->ids[k] + ": controlling element '" + id +
                          "' not found or has no branch current";
        delete C;
        return fail(KG_ELEMENT_NOT_FOUND, msg);
      }
      e.ctrl_branch = b->second;
    }
    C->el.push_back(e);
  }
  // Opt-in series resistance: after krets' unknowns and elements, so their
  // numbering is unchanged, each diode with rs > 0 gets an internal anode node
  // V(<id>:a) and a resistor <id>:rs from its anode to that node; the diode
  // then sits between the internal node and its cathode (SPICE's placement).
  if (flags & KG_NETLIST_DIODE_PARASITICS) {
    const size_t ne0 = C->el.size();
    for (size_t k = 0; k < ne0; k++) {
      if (C->recs[k].letter != 'D' || !(C->recs[k].p[7] > 0.0)) {
        if (C->recs[k].letter == 'D') C->recs[k].p[7] = 0.0;
        continue;
      }
      const double rs = C->recs[k].p[7];
      C->recs[k].p[7] = 0.0;
      const int32_t ia = (int32_t)C->names.size();
      C->names.push_back("V(" + C->ids[k] + ":a)");
      Rec rr;
      rr.letter = 'R';
      rr.name = C->recs[k].name + ":rs";
      rr.p[0] = rs;
      kg_element e{};
      e.type = KG_R;
      e.plus = C->el[k].plus;
      e.minus = ia;
      e.branch = -1;
      e.ctrl_plus = e.ctrl_minus = e.ctrl_branch = -1;
      C->el[k].plus = ia;
      C->recs.push_back(rr);
      C->ids.push_back(C->ids[k] + ":rs");
      C->el.push_back(e);
    }
  }
  *out = reinterpret_cast<kg_circuit*>(C);
  return KG_OK;
}

extern "C" void kg_circuit_free(kg_circuit* c) { delete reinterpret_cast<kg_circuit_impl*>(c); }

extern "C" int32_t kg_circuit_size(const kg_circuit* c, int32_t* n, int32_t* ne) {
  if (!c) return KG_INVALID_ARGUMENT;
  auto* C = reinterpret_cast<const kg_circuit_impl*>(c);
  if (n) *n = (int32_t)C->names.size();
  if (ne) *ne = (int32_t)C->el.size();
  return KG_OK;
}

extern "C" const char* kg_circuit_unknown_name(const kg_circuit* c, int32_t i) {
  auto* C = reinterpret_cast<const kg_circuit_impl*>(c);
  if (!C || i < 0 || i >= (int32_t)C->names.size()) return nullptr;
  return C->names[i].c_str();
}

extern "C" const char* kg_circuit_element_id(const kg_circuit* c, int32_t e) {
  auto* C = reinterpret_cast<const kg_circuit_impl*>(c);
  if (!C || e < 0 || e >= (int32_t)C->ids.size()) return nullptr;
  return C->ids[e].c_str();
}

extern "C" int32_t kg_circuit_tables(const kg_circuit* c, kg_element* el, double* params) {
  auto* C = reinterpret_cast<const kg_circuit_impl*>(c);
  if (!C) return KG_INVALID_ARGUMENT;
  for (size_t k = 0; k < C->el.size(); k++) {
    if (el) el[k] = C->el[k];
    if (params)
      for (int q = 0; q < KG_NPARAM; q++) params[k * KG_NPARAM + q] = C->recs[k].p[q];
  }
  return KG_OK;
}
