/* A C99 caller of the library's host-only entry points, compiled with
 * gcc -std=c99 -pedantic -Wall -Wextra -Werror against include/krets_gpu.h:
 * the header must be plain C for a Rust bindgen or any other C FFI. It reads
 * a netlist and a krets.toml text given on the command line and prints what
 * the library returns, one fact per line, for tests/test_host.py to compare
 * with the Python binding. No GPU call is made. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "krets_gpu.h"

static char* slurp(const char* path) {
  FILE* f = fopen(path, "rb");
  long n;
  char* buf;
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  n = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf = (char*)malloc((size_t)n + 1);
  if (buf && fread(buf, 1, (size_t)n, f) != (size_t)n) {
    free(buf);
    buf = NULL;
  }
  if (buf) buf[n] = '\0';
  fclose(f);
  return buf;
}

int main(int argc, char** argv) {
  char err[256];
  kg_circuit* c = NULL;
  kg_spec spec;
  kg_solver_config cfg;
  int32_t n = 0, ne = 0, i, r, layout[64];
  kg_element* el;
  double* prm;
  char* text;
  if (argc < 3) return 2;
  text = slurp(argv[1]);
  if (!text) return 2;
  r = kg_parse_netlist(text, &c, err, (int32_t)sizeof err);
  printf("parse %d\n", (int)r);
  if (r != KG_OK) return 1;
  kg_circuit_size(c, &n, &ne);
  printf("size %d %d\n", (int)n, (int)ne);
  for (i = 0; i < n; i++) printf("unknown %d %s\n", (int)i, kg_circuit_unknown_name(c, i));
  for (i = 0; i < ne; i++) printf("element %d %s\n", (int)i, kg_circuit_element_id(c, i));
  el = (kg_element*)malloc(sizeof(kg_element) * (size_t)ne);
  prm = (double*)malloc(sizeof(double) * (size_t)ne * KG_NPARAM);
  r = kg_circuit_tables(c, el, prm);
  printf("tables %d\n", (int)r);
  for (i = 0; i < ne; i++)
    printf("row %d %d %d %d %d %.17g\n", (int)i, (int)el[i].type, (int)el[i].plus, (int)el[i].minus,
           (int)el[i].branch, prm[(size_t)i * KG_NPARAM]);
  r = kg_plan_export(el, ne, n, 1, 1, prm, 6283.185307179586, NULL, 0, layout, 64);
  printf("plan %d %d %d\n", (int)(r > 0), (int)layout[0], (int)layout[1]);
  kg_default_config(&cfg);
  printf("config %.17g %.17g %.17g %d %d\n", cfg.reltol, cfg.iabstol, cfg.vabstol, (int)cfg.maxiter, (int)cfg.compat);
  kg_circuit_free(c);
  free(el);
  free(prm);
  free(text);
  text = slurp(argv[2]);
  if (!text) return 2;
  memset(&spec, 0, sizeof spec);
  r = kg_parse_spec(text, &spec, err, (int32_t)sizeof err);
  printf("spec %d %d %s\n", (int)r, (int)spec.kind, spec.circuit_path);
  free(text);
  return 0;
}
