/*
 * refcpu — CPU restatement of krets' MNA solve path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only the test suite,
 * the smoke check and bench.py's cpu_baseline leg may load it. The product
 * (libkrets_gpu.so) never links or calls it.
 *
 * Every unknown is numbered exactly as krets numbers it
 * (crates/krets-parser/src/parser.rs:105-125). Solutions are returned as flat
 * arrays in that numbering; complex results are interleaved (re, im).
 */
#ifndef KRETS_REFCPU_H
#define KRETS_REFCPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror krets_solver::Error (crates/krets-solver/src/error.rs:3-36). */
enum {
  REF_OK = 0,
  REF_MATRIX_BUILD = 1,
  REF_DECOMPOSITION = 2,
  REF_MAX_ITERATIONS = 3,
  REF_ELEMENT_NOT_FOUND = 4,
  REF_INVALID_ELEMENT_FORMAT = 5,
  REF_UNEXPECTED = 6,
  REF_PARSE = 7,
  REF_EMPTY_NETLIST = 8
};

/* Cost model of the solve loop. COMPAT reproduces krets' per-iteration work
 * (string-keyed maps, triplet vectors, hash-map summation, fresh ordering and
 * partial-pivot sparse LU every solve). LEAN computes the same numbers from
 * integer indices with the ordering reused. */
enum { REF_COST_COMPAT = 0, REF_COST_LEAN = 1 };

/* Transient methods. BE_COMPAT is krets' fixed-step backward Euler with every
 * quirk (source time lag, diode 0.5 V default, time key in convergence).
 * BE and TRAP are the clean integrators. */
enum { REF_BE_COMPAT = 0, REF_BE = 1, REF_TRAP = 2 };

typedef struct {
  double reltol;   /* SolverConfig::relative_tolerance (1e-3) */
  double iabstol;  /* current_absolute_tolerance (1e-12) */
  double vabstol;  /* voltage_absolute_tolerance (1e-6) */
  int32_t maxiter; /* maximum_iterations (300) */
  int32_t cost;    /* REF_COST_* */
} ref_config;

void ref_default_config(ref_config* cfg);

/* Parse a netlist (R/C/L/V/I/D subset). Returns NULL on error, with the
 * status in *status and a message in err. */
void* ref_parse(const char* netlist, int32_t* status, char* err, int32_t errlen);
/* ref_parse with options: flags & 1 reads the opt-in diode parasitics (series
 * resistance as an internal node V(<id>:a) plus a resistor <id>:rs numbered
 * after krets' own unknowns and elements; junction charge cjo vj m fc tt),
 * as the library's kg_parse_netlist_ext does. */
void* ref_parse_ext(const char* netlist, int32_t flags, int32_t* status, char* err, int32_t errlen);
void ref_free(void* circuit);
int32_t ref_num_unknowns(void* circuit);
int32_t ref_num_elements(void* circuit);
/* Name of unknown idx ("V(out)", "I(V1)"). */
const char* ref_unknown_name(void* circuit, int32_t idx);
/* Element description: type letter, node indices (-1 = ground), branch index,
 * primary value. Returns 0 on success. */
int32_t ref_element_info(void* circuit, int32_t e, char* type, int32_t* plus, int32_t* minus,
                         int32_t* branch, double* value);
/* Controlling node indices of E / G and controlling branch index of F / H
 * (-1 where not applicable). Returns 0 on success. */
int32_t ref_element_ctrl(void* circuit, int32_t e, int32_t* ctrl_plus, int32_t* ctrl_minus,
                         int32_t* ctrl_branch);
/* Overwrite an element's primary value (R, C, L, V dc, I, D Is, E/G/F/H gain). */
int32_t ref_set_value(void* circuit, int32_t e, double value);
const char* ref_last_error(void);

/* Analyses (krets semantics). x is [points][n] (real) or [points][n][2]. */
int32_t ref_op(void* c, const ref_config* cfg, double* x, int32_t* iters);
int32_t ref_dc(void* c, const ref_config* cfg, const char* element, double start, double stop,
               double step, double* x, int32_t maxpts, int32_t* npts);
/* AcAnalysis::generate_frequencies; variation 0 = Decade, 1 = Octave, 2 = Linear. */
int32_t ref_ac_frequencies(int32_t variation, int32_t points, double fstart, double fstop,
                           double* out, int32_t maxn);
int32_t ref_ac(void* c, const ref_config* cfg, const double* freqs, int32_t nf, double* x);
/* Fixed-step transient; x is [(nsteps+1)][n], nsteps = round(stop/h).
 * iters (optional) receives the Newton iteration count per step. */
int32_t ref_tran(void* c, const ref_config* cfg, double h, double stop, int32_t method,
                 double* x, int32_t maxpts, int32_t* npts, int32_t* iters);
/* Backward Euler with local-truncation-error step control (clean mode).
 * Writes accepted times and solutions; returns counts. */
int32_t ref_tran_lte(void* c, const ref_config* cfg, double tstop, double h0, double hmin,
                     double hmax, double lte_reltol, double lte_abstol, double* times,
                     double* x, int32_t maxpts, int32_t* npts, int32_t* newton_total,
                     int32_t* rejected);
/* Discrete adjoint of a fixed-step transient (BE or TRAP, linear circuits).
 * Objective J = sum_n w_n * (cvec . x_n), w = trapezoid-rule weights over
 * t_0..t_N. grad receives dJ/d(value) for every element (R, C, L; 0 for
 * others). */
int32_t ref_tran_adjoint(void* c, const ref_config* cfg, double h, int32_t nsteps,
                         int32_t method, const double* cvec, double* J, double* grad);

/* The LTE controller of ref_tran_lte with a record of every attempted step:
 * trace is [maxtr][5] = (t at the step's start, h, ratio = the largest
 * estimate / tolerance over the capacitors and inductors (-1 when the step has
 * no estimate: the first step, or a Newton failure), outcome (1 accepted,
 * 0 rejected by the estimate, -1 rejected after a Newton failure), the step
 * size the controller tries next). *ntr receives the number of rows. */
int32_t ref_tran_lte_trace(void* c, const ref_config* cfg, double tstop, double h0, double hmin,
                           double hmax, double lte_reltol, double lte_abstol, double* trace,
                           int32_t maxtr, int32_t* ntr);

/* Clean backward Euler on a caller's time grid (times[0] = 0, increasing):
 * the steps of ref_tran_lte without its controller, so that a run of the
 * library's LTE transient can be checked on its own accepted grid. x is
 * [npts][n]; x[0] is the clean operating point. */
int32_t ref_tran_grid(void* c, const ref_config* cfg, const double* times, int32_t npts, double* x,
                      int32_t* newton_total);

/* Single functions of the restatement, exposed so that tests can pin them on
 * their own (no krets entry point corresponds):
 * ref_wave_value: Pulse::value_at (kind 1, p = v1 v2 td tr tf pw per) and
 *   Sinusoidal::value_at (kind 2, p = vo va f td theta phase_deg)
 *   (voltage_source.rs:50-93); kind 0 returns p[0].
 * ref_parse_value: utils.rs:32-64 parse_value on one token; REF_OK with *v,
 *   or REF_PARSE.
 * ref_pnjlim: SPICE3 junction-voltage limiting of the clean Newton loops;
 *   *out receives the limited voltage; returns 1 if it limited, else 0. */
double ref_wave_value(int32_t kind, const double* p, double t);
int32_t ref_parse_value(const char* tok, double* v);
int32_t ref_pnjlim(double vnew, double vold, double vt, double vcrit, double* out);

/* Dense long-double LU with partial pivoting (n <= 256), real or complex.
 * A is row-major; complex arrays are interleaved. Returns 0 or 2 (singular). */
int32_t ref_dense_solve(int32_t n, const double* A, const double* b, double* x);
int32_t ref_dense_solve_c(int32_t n, const double* A, const double* b, double* x);
/* Assemble krets' MNA matrix and RHS densely for one analysis point, in
 * reference numbering: kind 0 = DC at guess x (NULL = empty map), 1 = AC at
 * frequency f with OP xop (complex A,b interleaved). */
int32_t ref_assemble_dc(void* c, const double* guess, double* A, double* b);
int32_t ref_assemble_ac(void* c, const double* xop, double f, double* A, double* b);

#ifdef __cplusplus
}
#endif
#endif
