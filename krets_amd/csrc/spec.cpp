// Reader of krets' analysis spec files (krets.toml): the TOML subset that
// `AnalysisSpec` deserialises (crates/krets-parser/src/analyses.rs:8-91):
//
//   circuit_path = "ac.cir"
//   [analysis.ac]                     (or .dc / .transient; `analysis = "op"`)
//   sweep = { variation = "Linear", total_points = 10 }
//   fstart = 0.1
//   fstop = 1000.0
//
// Keys follow serde's names: DcAnalysis {element, start, stop, step_size},
// AcAnalysis {sweep {variation: Decade|Octave|Linear, points_per_decade |
// points_per_octave | total_points}, fstart, fstop}, TransientAnalysis
// {time_step, stop_time}. Values: basic strings, integers, floats, inline
// tables; `#` comments.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "krets_gpu.h"

namespace {

struct Value {
  enum Kind { STR, NUM, TABLE } kind = NUM;
  std::string s;
  double num = 0.0;
  bool is_int = false;
  std::map<std::string, Value> table;
};

struct Parser {
  const std::string& t;
  size_t i = 0;
  int line = 1;
  std::string err;
  explicit Parser(const std::string& text) : t(text) {}

  void ws() {  // spaces and tabs (not newlines)
    while (i < t.size() && (t[i] == ' ' || t[i] == '\t' || t[i] == '\r')) i++;
  }
  void skip_comment() {
    if (i < t.size() && t[i] == '#')
      while (i < t.size() && t[i] != '\n') i++;
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = "krets.toml line " + std::to_string(line) + ": " + m;
    return false;
  }
  bool key(std::string& k) {
    ws();
    size_t b = i;
    if (i < t.size() && t[i] == '"') {
      Value v;
      if (!str(v)) return false;
      k = v.s;
      return true;
    }
    while (i < t.size() && (std::isalnum((unsigned char)t[i]) || t[i] == '_' || t[i] == '-')) i++;
    if (i == b) return fail("expected a key");
    k = t.substr(b, i - b);
    return true;
  }
  bool str(Value& v) {
    if (i >= t.size() || t[i] != '"') return fail("expected a string");
    i++;
    v.kind = Value::STR;
    while (i < t.size() && t[i] != '"') {
      if (t[i] == '\n') return fail("unterminated string");
      if (t[i] == '\\' && i + 1 < t.size()) {
        char e = t[i + 1];
        v.s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
        i += 2;
        continue;
      }
      v.s += t[i++];
    }
    if (i >= t.size()) return fail("unterminated string");
    i++;
    return true;
  }
  bool value(Value& v) {
    ws();
    if (i >= t.size()) return fail("expected a value");
    if (t[i] == '"') return str(v);
    if (t[i] == '{') {
      i++;
      v.kind = Value::TABLE;
      ws();
      if (i < t.size() && t[i] == '}') {
        i++;
        return true;
      }
      for (;;) {
        std::string k;
        if (!key(k)) return false;
        ws();
        if (i >= t.size() || t[i] != '=') return fail("expected '=' in inline table");
        i++;
        Value x;
        if (!value(x)) return false;
        v.table[k] = x;
        ws();
        if (i < t.size() && t[i] == ',') {
          i++;
          continue;
        }
        if (i < t.size() && t[i] == '}') {
          i++;
          return true;
        }
        return fail("expected ',' or '}' in inline table");
      }
    }
    // number: TOML integers and floats (underscores allowed between digits)
    size_t b = i;
    std::string num;
    while (i < t.size() && (std::isalnum((unsigned char)t[i]) || t[i] == '+' || t[i] == '-' || t[i] == '.' ||
                            t[i] == '_')) {
      if (t[i] != '_') num += t[i];
      i++;
    }
    if (i == b) return fail("expected a value");
    char* end = nullptr;
    v.kind = Value::NUM;
    v.num = std::strtod(num.c_str(), &end);
    if (!end || *end) return fail("invalid number '" + num + "'");
    v.is_int = num.find_first_of(".eEni") == std::string::npos;
    return true;
  }
  // document: top-level pairs and [a.b] tables -> flat map "a.b.key"
  bool document(std::map<std::string, Value>& out) {
    std::string prefix;
    while (i < t.size()) {
      ws();
      skip_comment();
      if (i >= t.size()) break;
      if (t[i] == '\n') {
        i++;
        line++;
        continue;
      }
      if (t[i] == '[') {
        i++;
        std::string name;
        for (;;) {
          std::string part;
          if (!key(part)) return false;
          name += name.empty() ? part : "." + part;
          ws();
          if (i < t.size() && t[i] == '.') {
            i++;
            continue;
          }
          break;
        }
        if (i >= t.size() || t[i] != ']') return fail("expected ']'");
        i++;
        prefix = name + ".";
        out[name].kind = Value::TABLE;  // an (empty) table is itself a value
      } else {
        std::string k;
        if (!key(k)) return false;
        ws();
        if (i >= t.size() || t[i] != '=') return fail("expected '='");
        i++;
        Value v;
        if (!value(v)) return false;
        if (out.count(prefix + k)) return fail("duplicate key '" + prefix + k + "'");
        out[prefix + k] = v;
      }
      ws();
      skip_comment();
      if (i < t.size() && t[i] != '\n') return fail("unexpected text after a value");
    }
    return true;
  }
};

int32_t spec_err(char* err, int32_t errlen, const std::string& m, int32_t code) {
  if (err && errlen > 0) std::snprintf(err, errlen, "%s", m.c_str());
  return code;
}

// Fills `out` from the analysis tables under `pre` ("analysis." inside an
// AnalysisSpec, "" for a bare Analysis document). Serde's externally tagged
// enum (analyses.rs:27-41): exactly one variant, lowercase names.
int32_t read_analysis(const std::map<std::string, Value>& kv, const std::string& pre, kg_spec* out, char* err,
                      int32_t errlen) {
  auto num = [&](const std::string& k, double* d) -> bool {
    auto f = kv.find(pre + k);
    if (f == kv.end() || f->second.kind != Value::NUM) return false;
    *d = f->second.num;
    return true;
  };
  int kinds = 0;
  if (!pre.empty()) {  // the unit variant Op is the string "op"
    auto a = kv.find(pre.substr(0, pre.size() - 1));
    if (a != kv.end() && a->second.kind == Value::STR) {
      if (a->second.s != "op") return spec_err(err, errlen, "unknown analysis '" + a->second.s + "'", KG_PARSE);
      out->kind = KG_SPEC_OP;
      kinds++;
    }
  }
  for (auto& e : kv) {  // unknown variants
    if (e.first.compare(0, pre.size(), pre) != 0 || e.first.size() == pre.size()) continue;
    std::string v = e.first.substr(pre.size());
    v = v.substr(0, v.find('.'));
    if (v != "dc" && v != "ac" && v != "transient")
      return spec_err(err, errlen, "unknown analysis '" + v + "'", KG_PARSE);
  }
  if (kv.count(pre + "dc")) {
    kinds++;
    out->kind = KG_SPEC_DC;
    auto e = kv.find(pre + "dc.element");
    if (e == kv.end() || e->second.kind != Value::STR || e->second.s.size() >= sizeof(out->element))
      return spec_err(err, errlen, "dc: missing element", KG_PARSE);
    std::snprintf(out->element, sizeof(out->element), "%s", e->second.s.c_str());
    if (!num("dc.start", &out->start) || !num("dc.stop", &out->stop) || !num("dc.step_size", &out->step_size))
      return spec_err(err, errlen, "dc: start, stop and step_size are required", KG_PARSE);
  }
  if (kv.count(pre + "ac")) {
    kinds++;
    out->kind = KG_SPEC_AC;
    auto s = kv.find(pre + "ac.sweep");
    if (s == kv.end() || s->second.kind != Value::TABLE)
      return spec_err(err, errlen, "ac: missing sweep table", KG_PARSE);
    auto& st = s->second.table;
    auto var = st.find("variation");
    if (var == st.end() || var->second.kind != Value::STR)
      return spec_err(err, errlen, "ac.sweep: missing variation", KG_PARSE);
    // AcSweep is tagged by "variation" (analyses.rs:56-65)
    const char* field = nullptr;
    if (var->second.s == "Decade") out->sweep = KG_SWEEP_DECADE, field = "points_per_decade";
    else if (var->second.s == "Octave") out->sweep = KG_SWEEP_OCTAVE, field = "points_per_octave";
    else if (var->second.s == "Linear") out->sweep = KG_SWEEP_LINEAR, field = "total_points";
    else return spec_err(err, errlen, "ac.sweep: unknown variation '" + var->second.s + "'", KG_PARSE);
    auto pts = st.find(field);
    if (pts == st.end() || pts->second.kind != Value::NUM || !pts->second.is_int || pts->second.num < 0 ||
        pts->second.num > 4294967295.0)
      return spec_err(err, errlen, std::string("ac.sweep: ") + field + " must be a u32", KG_PARSE);
    out->points = (int64_t)pts->second.num;
    if (!num("ac.fstart", &out->fstart) || !num("ac.fstop", &out->fstop))
      return spec_err(err, errlen, "ac: fstart and fstop are required", KG_PARSE);
  }
  if (kv.count(pre + "transient")) {
    kinds++;
    out->kind = KG_SPEC_TRANSIENT;
    if (!num("transient.time_step", &out->time_step) || !num("transient.stop_time", &out->stop_time))
      return spec_err(err, errlen, "transient: time_step and stop_time are required", KG_PARSE);
  }
  if (kinds != 1) return spec_err(err, errlen, "exactly one analysis is required", KG_PARSE);
  return KG_OK;
}

int32_t parse(const char* text, kg_spec* out, char* err, int32_t errlen, bool spec) {
  if (!text || !out) return spec_err(err, errlen, "null argument", KG_INVALID_ARGUMENT);
  std::memset(out, 0, sizeof(*out));
  if (err && errlen > 0) err[0] = 0;
  std::string t(text);
  Parser p(t);
  std::map<std::string, Value> kv;
  if (!p.document(kv)) return spec_err(err, errlen, p.err, KG_PARSE);
  if (!spec) return read_analysis(kv, "", out, err, errlen);
  auto it = kv.find("circuit_path");
  if (it == kv.end() || it->second.kind != Value::STR)
    return spec_err(err, errlen, "missing circuit_path", KG_PARSE);
  if (it->second.s.size() >= sizeof(out->circuit_path))
    return spec_err(err, errlen, "circuit_path too long", KG_PARSE);
  std::snprintf(out->circuit_path, sizeof(out->circuit_path), "%s", it->second.s.c_str());
  return read_analysis(kv, "analysis.", out, err, errlen);
}

}  // namespace

extern "C" int32_t kg_parse_spec(const char* text, kg_spec* out, char* err, int32_t errlen) {
  return parse(text, out, err, errlen, true);
}

extern "C" int32_t kg_parse_analysis(const char* text, kg_spec* out, char* err, int32_t errlen) {
  return parse(text, out, err, errlen, false);
}
