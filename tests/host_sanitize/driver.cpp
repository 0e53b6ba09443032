// Host-side robustness driver, built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_host_sanitize.py). It feeds the
// library's untrusted-input code, the netlist reader, the krets.toml reader,
// the symbolic pass and the run-time source generators, with krets' own
// input files and with seeded mutations of them (truncations, byte
// substitutions, duplicated and dropped lines), and runs every plan and
// generator the library would run on what parses. Any out-of-bounds access,
// use after free, leak or undefined behaviour aborts the process. No GPU
// call is made.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "kg_jit.h"
#include "krets_gpu.h"

namespace kg {
std::vector<int32_t> build_pattern(const kg_element* el, int ne, int n, AnalysisKind ak, Ordering ord,
                                   const double* rep, double s_rep, PatLayout& L, std::string& err);
bool build_lin_plan(const PatLayout& L, const std::vector<int32_t>& b, std::vector<uint16_t>& out);
}  // namespace kg

namespace {

long g_parsed = 0, g_rejected = 0, g_plans = 0, g_sources = 0;

void plans(const std::vector<kg_element>& el, const std::vector<double>& prm, int n) {
  const int ne = (int)el.size();
  for (int ak = 0; ak < 3; ak++)
    for (int ord = 0; ord < 4; ord++) {
      kg::PatLayout L{};
      std::string err;
      const double s = ak == 1 ? 2 * 3.141592653589793 * 1e3 : ak == 2 ? 1e9 : 0.0;
      std::vector<int32_t> b = kg::build_pattern(el.data(), ne, n, (kg::AnalysisKind)ak, (kg::Ordering)ord,
                                                 prm.data(), s, L, err);
      if (L.status != KG_OK) continue;
      g_plans++;
      if (ak == 2) {
        std::vector<uint16_t> lin;
        kg::build_lin_plan(L, b, lin);
      }
      if (ord != 1) continue;  // the generators work on minimum-degree plans
      std::string src, why;
      if (ak == 1)
        for (int m = 0; m < 3; m++) g_sources += kg::jit_ac_source(L, b.data(), m > 0, m == 2, src, why);
      if (ak == 0) g_sources += kg::jit_dc_source(L, b.data(), src, why);
      if (ak == 2) {
        g_sources += kg::jit_tran_source(L, b.data(), src, why);
        g_sources += kg::jit_lte_source(L, b.data(), src, why);
      }
    }
}

void netlist(const std::string& text) {
  kg_circuit* c = nullptr;
  char err[256];
  if (kg_parse_netlist(text.c_str(), &c, err, sizeof err) != KG_OK) {
    g_rejected++;
    return;
  }
  g_parsed++;
  int32_t n = 0, ne = 0;
  kg_circuit_size(c, &n, &ne);
  for (int i = -1; i <= n; i++) (void)kg_circuit_unknown_name(c, i);
  for (int e = -1; e <= ne; e++) (void)kg_circuit_element_id(c, e);
  std::vector<kg_element> el(ne > 0 ? ne : 1);
  std::vector<double> prm((size_t)(ne > 0 ? ne : 1) * KG_NPARAM);
  if (kg_circuit_tables(c, el.data(), prm.data()) == KG_OK && ne > 0 && n > 0 && n <= 4000) {
    el.resize(ne);
    prm.resize((size_t)ne * KG_NPARAM);
    plans(el, prm, n);
  }
  kg_circuit_free(c);
}

void spec(const std::string& text) {
  kg_spec s;
  char err[256];
  if (kg_parse_spec(text.c_str(), &s, err, sizeof err) == KG_OK) g_parsed++;
  else g_rejected++;
  if (kg_parse_analysis(text.c_str(), &s, err, sizeof err) == KG_OK) g_parsed++;
  else g_rejected++;
}

std::string mutate(const std::string& s, std::mt19937_64& rng) {
  std::string t = s;
  const char alphabet[] = "0123456789.eE+-kKmMuUnNpPfFgGtT()=_ \t\n*%.$XVRLCIDEGFHQ";
  switch (rng() % 5) {
    case 0:  // truncate
      t.resize(t.empty() ? 0 : rng() % t.size());
      break;
    case 1:  // substitute bytes
      for (int k = 0, m = 1 + (int)(rng() % 4); k < m && !t.empty(); k++)
        t[rng() % t.size()] = alphabet[rng() % (sizeof alphabet - 1)];
      break;
    case 2: {  // duplicate a line
      std::vector<std::string> lines;
      std::istringstream in(t);
      for (std::string l; std::getline(in, l);) lines.push_back(l);
      if (lines.empty()) break;
      lines.insert(lines.begin() + rng() % lines.size(), lines[rng() % lines.size()]);
      t.clear();
      for (auto& l : lines) t += l + "\n";
      break;
    }
    case 3: {  // drop a line
      std::vector<std::string> lines;
      std::istringstream in(t);
      for (std::string l; std::getline(in, l);) lines.push_back(l);
      if (lines.empty()) break;
      lines.erase(lines.begin() + rng() % lines.size());
      t.clear();
      for (auto& l : lines) t += l + "\n";
      break;
    }
    default:  // insert bytes
      for (int k = 0, m = 1 + (int)(rng() % 6); k < m; k++)
        t.insert(t.empty() ? 0 : rng() % t.size(), 1, alphabet[rng() % (sizeof alphabet - 1)]);
  }
  return t;
}

}  // namespace

// usage: driver <mutations per file> <seed> file...  (*.toml: spec, else netlist)
int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int nm = std::atoi(argv[1]);
  std::mt19937_64 rng(std::strtoull(argv[2], nullptr, 10));
  for (int a = 3; a < argc; a++) {
    std::ifstream f(argv[a]);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const bool is_spec = std::strstr(argv[a], ".toml") != nullptr;
    (is_spec ? spec : netlist)(text);
    for (int k = 0; k < nm; k++) (is_spec ? spec : netlist)(mutate(text, rng));
  }
  std::printf("parsed %ld rejected %ld plans %ld sources %ld\n", g_parsed, g_rejected, g_plans, g_sources);
  return 0;
}
