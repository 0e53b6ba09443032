/*
 * krets_gpu.h — C ABI of the MI355X (gfx950) MNA solve backend for krets.
 *
 * This library replaces one path of krets: MNA stamping into a fixed sparse
 * pattern, sparse LU, triangular solves, and the Newton / DC-sweep / AC-sweep /
 * transient loops around them. It keeps krets' unknown numbering
 * (crates/krets-parser/src/parser.rs:105-125), sign conventions and status
 * codes, so results map back 1:1 onto krets' AnalysisResult maps.
 *
 * All functions are synchronous, return a kg_status, and never throw or abort.
 * A context owns its device memory and one HIP stream; it is not re-entrant
 * (one context per host thread), like krets' `Solver::solve(&mut self, ..)`.
 * Output buffers are caller-owned, either host memory (KG_MEM_HOST) or device
 * memory on the context's GPU (KG_MEM_DEVICE). The per-instance status array
 * and error estimates of kg_op / kg_dc_sweep / kg_ac_sweep / kg_transient /
 * kg_transient_lte may be either, whatever `mem` says of the other outputs
 * (the library checks the pointer): a device solution with host statuses is
 * the common case.
 */
#ifndef KRETS_GPU_H
#define KRETS_GPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 0..6 mirror krets_solver::Error
 * (crates/krets-solver/src/error.rs:3-36): MatrixBuild, DecompositionFailed /
 * MatrixDecomposition, MaximumIterationsExceeded, ElementNotFound,
 * InvalidElementFormat, Unexpected. 7..8 mirror krets_parser errors
 * (ParseError, EmptyNetlist). */
typedef enum {
  KG_OK = 0,
  KG_MATRIX_BUILD = 1,
  KG_DECOMPOSITION = 2,
  KG_MAX_ITERATIONS = 3,
  KG_ELEMENT_NOT_FOUND = 4,
  KG_INVALID_ELEMENT_FORMAT = 5,
  KG_UNEXPECTED = 6,
  KG_PARSE = 7,
  KG_EMPTY_NETLIST = 8,
  KG_INVALID_ARGUMENT = 9,
  KG_HIP_ERROR = 100
} kg_status;

/* KG_E .. KG_H are the linear controlled sources of krets' README to-do list
 * (README.md:56-62), which krets does not implement yet: E (VCVS), G (VCCS),
 * F (CCCS), H (CCVS), with SPICE syntax and sign conventions. */
typedef enum {
  KG_R = 0, KG_C = 1, KG_L = 2, KG_V = 3, KG_I = 4, KG_D = 5,
  KG_E = 6, KG_G = 7, KG_F = 8, KG_H = 9
} kg_element_type;
typedef enum { KG_WAVE_NONE = 0, KG_WAVE_PULSE = 1, KG_WAVE_SIN = 2 } kg_wave_kind;
typedef enum { KG_FP64 = 0, KG_MIXED = 1 } kg_precision;
typedef enum { KG_BE_COMPAT = 0, KG_BE = 1, KG_TRAP = 2 } kg_method;
typedef enum { KG_MEM_HOST = 0, KG_MEM_DEVICE = 1 } kg_mem;
/* Work mapping: threads per system. 0 = autotuned. */
typedef enum { KG_MAP_AUTO = 0 } kg_mapping;

/* One element in krets' numbering: plus/minus are the indices of V(plus) and
 * V(minus) (-1 for ground "0"), branch the index of I(<id>) or -1. E and H
 * own a branch current, like V. ctrl_plus / ctrl_minus are the controlling
 * node indices of E and G (-1 for ground); ctrl_branch is the index of the
 * controlling element's branch current for F and H. The ctrl fields are -1
 * for every other type. */
typedef struct {
  int32_t type;
  int32_t plus;
  int32_t minus;
  int32_t branch;
  int32_t g2; /* Group-2 flag of R/C (capacitor.rs:172) */
  int32_t ctrl_plus;
  int32_t ctrl_minus;
  int32_t ctrl_branch;
} kg_element;

/* Per (instance, element) parameter record of KG_NPARAM doubles:
 *   R: [0] ohms          C: [0] farads        L: [0] henries
 *   I: [0] amps          D: [0] Is, [1] N, [2] CJO, [3] VJ, [4] M, [5] FC,
 *                           [6] TT (junction charge; CJO = TT = 0: none)
 *   E: [0] voltage gain  G: [0] transconductance (S)
 *   F: [0] current gain  H: [0] transresistance (ohms)
 *   V: [0] dc value, [1] ac magnitude, [2] waveform kind (kg_wave_kind),
 *      [3..9] PULSE(v1 v2 td tr tf pw per) or SIN(vo va f td theta phase). */
#define KG_NPARAM 10

/* A batch of circuit instances. With shared_topology = 1 every instance uses
 * elements[0..n_elements) and n_unknowns[0]; params is
 * [n_instances][n_elements][KG_NPARAM]. Otherwise instance i uses
 * elements[elem_offset[i]..elem_offset[i+1]) and n_unknowns[i], and its
 * parameter records start at params + elem_offset[i] * KG_NPARAM. */
typedef struct {
  int32_t n_instances;
  int32_t shared_topology;
  int32_t n_elements;
  int32_t reserved;
  const int32_t* elem_offset;
  const int32_t* n_unknowns;
  const kg_element* elements;
  const double* params;
} kg_batch_desc;

/* SolverConfig (crates/krets-solver/src/config.rs:4-36) plus mode flags.
 * compat = 1 reproduces krets' Newton semantics exactly (empty first map,
 * 0.5 V diode default, bookkeeping key in the convergence test, unlimited
 * diode exponent); compat = 0 is the clean mode (zero start, pnjlim). */
typedef struct {
  double reltol;
  double iabstol;
  double vabstol;
  int32_t maxiter;
  int32_t compat;
  /* Work mapping; 0 = autotuned. 1, 4, 16, 64, 256: threads per system of the
   * level-scheduled kernels. AC sweeps of a shared topology also take
   * 1000 + t (t = 1, 4, 16, 64: the plan and each instance's element records
   * staged in LDS by a persistent grid), 2001 (one lane per system, code
   * specialised to the topology at run time) and, for KG_MIXED, 2002 (as
   * 2001 with the solution split into a complex64 first solution in LDS and
   * a complex64 sum of corrections in registers); 2011 / 2012 are 2001 / 2002
   * on a balanced elimination order. Fixed-step transients also
   * take 1064 and 1256 (64 or 256 threads per system with the plan copied
   * into LDS once per launch) and, for FP64 on circuits without diodes,
   * 3000 + t, t = 64, 128, 256, 512, 1024 (one system per block of t threads with
   * the whole step loop in LDS: compact 16-bit plan, factors, solutions and
   * companion states). For FP64 on a shared topology, fixed-step and LTE
   * transients, operating points and DC sweeps (linear circuits) also take
   * 2001: one lane per instance, code specialised to the topology at run
   * time (registers only; small circuits). */
  int32_t threads_per_system;
  /* KG_MIXED: refinement steps. 0 = default: 2, except that the lane AC
   * kernels (2002 / 2012) stop after the first pass where every lane of a
   * wave predicts an error below 1e-11. */
  int32_t refine_iters;
} kg_solver_config;

typedef struct kg_ctx kg_ctx;
typedef struct kg_circuit kg_circuit;

void kg_default_config(kg_solver_config* cfg);

/* Context lifetime. */
int32_t kg_create(kg_ctx** ctx, int32_t device);
int32_t kg_destroy(kg_ctx* ctx);
const char* kg_last_error(const kg_ctx* ctx);

/* Replaces faer's per-solve symbolic analysis (called at solver/op.rs:48-51,
 * dc.rs:81-84, ac.rs:67-70, transient.rs:74-77) with one host symbolic
 * pass per distinct topology: pattern, static pivot order (structural
 * matching plus fill-reducing order), fill, level sets; then uploads the
 * batch. The pivots come from the topology's first instance; an instance
 * whose static-pivot factorisation is not backward stable is reported, not
 * returned silently: operating points, DC sweeps and transients (FP64) bound
 * each returned solution's normwise backward error by the LU growth of its
 * factors and report KG_DECOMPOSITION above 2^-30 (krets' partial pivoting
 * would have re-pivoted); AC sweeps re-solve such systems with partial
 * pivoting (kg_ac_sweep). */
int32_t kg_load_batch(kg_ctx* ctx, const kg_batch_desc* batch);

/* op::solve (crates/krets-solver/src/solver/op.rs:15-86) for every instance.
 * x: [n_instances][n] doubles. status/iters: [n_instances] (may be NULL).
 * Returns the first failing instance's status (KG_OK if all succeeded). */
int32_t kg_op(kg_ctx* ctx, const kg_solver_config* cfg, double* x, int32_t* status, int32_t* iters,
              int32_t mem);

/* dc::solve (solver/dc.rs:14-124): sweeps element `element` (index within the
 * instance's element list; must be V or I) over start + i*step,
 * i < |(stop-start)/step| + 1. x: [n_instances][points][n]. */
int32_t kg_dc_sweep(kg_ctx* ctx, const kg_solver_config* cfg, int32_t element, double start, double stop,
                    double step, double* x, int32_t max_points, int32_t* n_points, int32_t* status,
                    int32_t* iters, int32_t mem);

/* ac::solve (solver/ac.rs:17-109) over an explicit frequency list (positive
 * entries; krets skips f <= 0). x: [n_instances][nf][n] complex128
 * interleaved (re, im). With KG_MIXED, the factorisation runs in complex64
 * and FP64 iterative refinement restores accuracy; err_est[instance] (may be
 * NULL) receives the largest normwise relative size of the last refinement
 * correction. With KG_FP64, err_est receives a bound on the normwise
 * backward error |b - A x| / (|A| |x| + |b|) of every solve of the instance.
 * Pivots are fixed per topology; every solve checks its componentwise
 * backward error in FP64 (mixed: its refinement estimate), and solves that
 * fail are solved again with partial pivoting at their own frequency, as
 * krets' LU does (systems of up to 1024 unknowns; larger ones, and systems
 * singular under partial pivoting, report KG_DECOMPOSITION). */
int32_t kg_ac_sweep(kg_ctx* ctx, const kg_solver_config* cfg, const double* freqs, int32_t nf,
                    int32_t precision, double* x, int32_t* status, double* err_est, int32_t mem);

/* Allocation-free AC sweep. No counterpart in krets: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation). It computes exactly what
 * kg_ac_sweep computes with mem = KG_MEM_DEVICE, through the same kernels,
 * but calls no hipMalloc, hipFree or other allocating routine: the
 * intermediates (operating point and its offsets, output offsets, the copy
 * of freqs, the accuracy guard's queue) live in a static device array of
 * 8 MiB per GPU that the library holds until the process ends.
 * x ([n_instances][nf][n] complex128), status [n_instances] and err_est
 * [n_instances] (may be NULL) must be device memory of the context's GPU.
 * Preconditions: the batch is loaded, and kg_ac_sweep has run on this
 * context with the same precision, frequency range and solver settings, so
 * that plans, compiled kernels and the work mapping are cached; otherwise
 * KG_INVALID_ARGUMENT. A workload whose intermediates exceed the array is
 * rejected with KG_INVALID_ARGUMENT and a message giving both sizes. Every
 * call on one GPU, from any context, shares that GPU's array, which keeps the
 * last call's contents: calls that overlap on one GPU race. Returns the first
 * failing instance's status when the context's pinned staging area (set up
 * by the earlier allocating call) can hold the statuses, else KG_OK once the
 * kernels have completed; the per-instance statuses are in `status`. */
int32_t kg_ac_sweep_noalloc(kg_ctx* ctx, const kg_solver_config* cfg, const double* freqs, int32_t nf,
                            int32_t precision, double* x, int32_t* status, double* err_est);

/* transient::solve (solver/transient.rs:14-114) with a fixed step h and
 * nsteps steps (krets: nsteps = round(stop/h)). method KG_BE_COMPAT is
 * krets' integrator with its quirks; KG_BE and KG_TRAP are clean.
 * x: [n_instances][nsteps+1][n] (row 0 = operating point); iters:
 * [n_instances][nsteps+1] Newton iterations (may be NULL). One persistent
 * launch runs all steps of all instances. err_est [n_instances] (may be
 * NULL): with KG_MIXED the refinement's estimate of the relative error,
 * with KG_FP64 the largest bound on the normwise backward error of the
 * instance's solves (the static-pivot guard of kg_load_batch). */
int32_t kg_transient(kg_ctx* ctx, const kg_solver_config* cfg, double h, int32_t nsteps, int32_t method,
                     int32_t precision, double* x, int32_t* status, int32_t* iters, double* err_est,
                     int32_t mem);

/* Backward Euler with local-truncation-error step control (no krets
 * counterpart; krets is fixed-step only). Accepted points per instance are
 * written to t: [n_instances][max_points] and x: [n_instances][max_points][n].
 * n_points, newton_total, rejected: [n_instances]. */
int32_t kg_transient_lte(kg_ctx* ctx, const kg_solver_config* cfg, double tstop, double h0, double hmin,
                         double hmax, double lte_reltol, double lte_abstol, int32_t max_points, double* t,
                         double* x, int32_t* n_points, int32_t* newton_total, int32_t* rejected,
                         int32_t* status, int32_t mem);

/* Discrete adjoint of a fixed-step linear transient (no krets counterpart).
 * Objective J = sum_n w_n * (cvec . x_n) with trapezoid weights w over
 * t_0..t_nsteps; cvec: [n] in krets numbering, shared by all instances.
 * Reuses the stored LU factors for transpose solves and recomputes states
 * from checkpoints taken every `checkpoint_every` steps.
 * J: [n_instances]; grad: [n_instances][n_elements] = dJ/d(param[0]) of every
 * R, C and L (0 for other elements). */
int32_t kg_adjoint_transient(kg_ctx* ctx, const kg_solver_config* cfg, double h, int32_t nsteps,
                             int32_t method, int32_t precision, const double* cvec, int32_t checkpoint_every,
                             double* J, double* grad, int32_t* status, int32_t mem);
/* The two halves of kg_adjoint_transient (no krets counterpart): the forward
 * pass returns J and keeps the checkpoints in the context; the backward pass
 * returns grad for the most recent forward pass. Batches with a shared
 * topology and linear circuits (R, C, L, V, I); method KG_BE or KG_TRAP.
 * threads_per_system: 1..256 (level-scheduled kernels), 3064 / 3256 / 3512
 * (FP64, the step loops in LDS, one system per block of 64 / 256 / 512
 * threads), 0 = the fastest forward pass among them, timed once per
 * workload. */
int32_t kg_adjoint_forward(kg_ctx* ctx, const kg_solver_config* cfg, double h, int32_t nsteps, int32_t method,
                           int32_t precision, const double* cvec, int32_t checkpoint_every, double* J,
                           int32_t* status, int32_t mem);
int32_t kg_adjoint_backward(kg_ctx* ctx, double* grad, int32_t* status, int32_t mem);

/* Graph replay (no krets counterpart; on by default): a call whose work
 * mapping is already known runs its launch sequence as a captured hipGraph,
 * captured once per distinct argument set and replayed afterwards. Calls made
 * while kernel timing is on run eagerly. Applies to kg_ac_sweep. */
int32_t kg_use_graphs(kg_ctx* ctx, int32_t enable);

/* Kernel timing (no krets counterpart): HIP events on the context's stream
 * time every analysis kernel launched while timing is on, per kernel. Reports
 * the dominant kernel since the previous call (the one with the largest
 * accumulated time): its name, accumulated ms and launch count; then resets
 * and turns timing on (enable = 1) or off. */
int32_t kg_profile(kg_ctx* ctx, int32_t enable, double* total_ms, int64_t* launches, const char** kernel);
/* Every kernel timed since the previous kg_profile call, one line each,
 * "<name> <accumulated ms> <launches>\n", into buf (NUL-terminated,
 * truncated to cap); no reset. Returns the length it needed. */
int32_t kg_profile_summary(const kg_ctx* ctx, char* buf, int32_t cap);

/* Autotuner report (no krets counterpart): the candidate mappings (threads
 * per system) timed by the most recent autotuning run and their kernel times
 * in ms. Returns the number of candidates (copying at most cap of them). The
 * winner is cached per workload in the context, and in the JSON file named by
 * the environment variable KRETS_GPU_TUNE_FILE when it is set. */
int32_t kg_last_autotune(const kg_ctx* ctx, int32_t* tps, float* ms, int32_t cap);

/* Diagnostics (no krets counterpart): the symbolic plan the library would
 * build for one topology. kind: 0 DC, 1 AC, 2 transient, 3 the compact 16-bit
 * plan of the LDS-resident linear transient kernel (uint16 words, two per
 * int32; -KG_UNEXPECTED if the topology does not qualify); ordering: 0 natural,
 * 1 minimum degree, 2 nested dissection, 3 balanced (minimum fill, ties to
 * the shallowest elimination-tree position). params (may be NULL): one instance's
 * parameter records, used with s_rep (alpha/h or 2*pi*f) for the static pivot
 * choice. Returns the plan length in int32 (copying at most cap entries to
 * buf) or -status. layout receives the plan's header (PatLayout, see
 * krets_amd/csrc/kg_internal.h). */
int32_t kg_plan_export(const kg_element* elements, int32_t n_elements, int32_t n_unknowns, int32_t kind,
                       int32_t ordering, const double* params, double s_rep, int32_t* buf, int32_t cap,
                       int32_t* layout, int32_t layout_cap);

/* Diagnostic: the HIP source of the run-time specialised lane-per-system AC
 * kernel for one topology (minimum-degree plan, static pivots from params at
 * reactive scale s_rep; refine_iters < 0 selects the mixed precision form
 * with the solution split into a complex64 first solution in LDS and a
 * complex64 sum of corrections in registers; the refinement count is a
 * launch argument). The mixed precision forms read element values from the
 * value table blocked by 64 instances (entry (i / 64 * ne + e) * 64 + i % 64)
 * and run 64 consecutive instances at one frequency per wave. No
 * counterpart in krets. Returns the source length
 * (copying at most cap bytes, NUL-terminated, to buf) or -status, with the
 * reason the topology is not specialised in buf. */
int32_t kg_jit_ac_source(const kg_element* elements, int32_t n_elements, int32_t n_unknowns, const double* params,
                         double s_rep, int32_t precision, int32_t refine_iters, char* buf, int32_t cap);

/* Diagnostic: the HIP source of the run-time specialised lane-per-instance
 * kernel of kg_op / kg_dc_sweep for linear circuits (minimum-degree plan of
 * the DC stamps, static pivots from params). No counterpart in krets.
 * Returns the source length (copying at most cap bytes, NUL-terminated, to
 * buf) or -status, with the reason in buf. */
int32_t kg_jit_dc_source(const kg_element* elements, int32_t n_elements, int32_t n_unknowns, const double* params,
                         char* buf, int32_t cap);

/* Diagnostic: the HIP source of the run-time specialised lane-per-instance
 * kernel of kg_transient (FP64; minimum-degree plan of the transient stamps,
 * static pivots from params at step h with the method's companion scale; the
 * method itself is a launch argument). No counterpart in krets. Returns the
 * source length (copying at most cap bytes, NUL-terminated, to buf) or
 * -status, with the reason in buf. */
int32_t kg_jit_tran_source(const kg_element* elements, int32_t n_elements, int32_t n_unknowns, const double* params,
                           double h, int32_t method, char* buf, int32_t cap);

/* Diagnostic: the HIP source of the run-time specialised lane-per-instance
 * kernel of kg_transient_lte for one topology (minimum-degree plan of the
 * transient stamps, static pivots from params at step h0). No counterpart in
 * krets. Returns the source length (copying at most cap bytes,
 * NUL-terminated, to buf) or -status, with the reason in buf. */
int32_t kg_jit_lte_source(const kg_element* elements, int32_t n_elements, int32_t n_unknowns, const double* params,
                          double h0, char* buf, int32_t cap);

/* Netlist reader for the R/L/C/V/I/D subset of krets' grammar
 * (crates/krets-parser/src/parser.rs:27-154). */
int32_t kg_parse_netlist(const char* text, kg_circuit** out, char* err, int32_t errlen);

/* kg_parse_netlist with options (no krets counterpart). flags = 0 reads
 * exactly as kg_parse_netlist. KG_NETLIST_DIODE_PARASITICS models what krets
 * parses and ignores (models/diode.rs:14-37) or does not parse at all:
 *  - rs > 0: an internal anode node V(<id>:a) and a resistor <id>:rs from
 *    the anode to it, both numbered after every krets unknown and element,
 *    so krets' own numbering is unchanged;
 *  - cjo (or cj0), vj, m, fc, tt: SPICE junction charge, into the diode's
 *    parameter record [2..6] (defaults cjo 0, vj 1, m 0.5, fc 0.5, tt 0),
 *    used by AC sweeps (small-signal capacitance at the operating point) and
 *    by the clean transients KG_BE / KG_TRAP and kg_transient_lte (charge
 *    companion); operating points, DC sweeps and KG_BE_COMPAT ignore it, as
 *    krets' integrator has no junction charge. */
#define KG_NETLIST_DIODE_PARASITICS 1
int32_t kg_parse_netlist_ext(const char* text, int32_t flags, kg_circuit** out, char* err, int32_t errlen);
void kg_circuit_free(kg_circuit* c);
int32_t kg_circuit_size(const kg_circuit* c, int32_t* n_unknowns, int32_t* n_elements);
/* Name of unknown idx, e.g. "V(out)" or "I(V1)"; NULL if out of range. */
const char* kg_circuit_unknown_name(const kg_circuit* c, int32_t idx);
/* Identifier of element e, e.g. "V1" (Element::identifier, elements.rs:314). */
const char* kg_circuit_element_id(const kg_circuit* c, int32_t e);
/* Fills elements[n_elements] and params[n_elements][KG_NPARAM]. */
int32_t kg_circuit_tables(const kg_circuit* c, kg_element* elements, double* params);

/* Analysis spec (krets.toml) in the shape of krets_parser::analyses::
 * AnalysisSpec (crates/krets-parser/src/analyses.rs:8-76): one analysis,
 * selected by kind; fields of the other kinds stay zero. */
typedef enum { KG_SPEC_OP = 0, KG_SPEC_DC = 1, KG_SPEC_AC = 2, KG_SPEC_TRANSIENT = 3 } kg_spec_kind;
typedef enum { KG_SWEEP_DECADE = 0, KG_SWEEP_OCTAVE = 1, KG_SWEEP_LINEAR = 2 } kg_sweep_kind;
typedef struct {
  int32_t kind;              /* kg_spec_kind */
  char circuit_path[1024];   /* as written; krets-cli resolves it against the spec's folder */
  char element[64];          /* DC: swept source (DcAnalysis::element) */
  double start, stop, step_size;
  int32_t sweep;             /* AC: kg_sweep_kind (AcSweep's "variation" tag) */
  int32_t reserved;
  int64_t points;            /* AC: points_per_decade / points_per_octave / total_points (u32) */
  double fstart, fstop;
  double time_step, stop_time;
} kg_spec;

/* Reads a krets.toml text, as AnalysisSpec::from_file deserialises it
 * (crates/krets-parser/src/analyses.rs:19-23). Returns KG_OK or KG_PARSE with
 * a message in err. */
int32_t kg_parse_spec(const char* toml_text, kg_spec* out, char* err, int32_t errlen);
/* Reads a bare Analysis document ("[transient]\ntime_step = ..."), as
 * toml::from_str::<Analysis> does in krets' own parser tests
 * (crates/krets-parser/src/analyses.rs:170-186); circuit_path stays empty. */
int32_t kg_parse_analysis(const char* toml_text, kg_spec* out, char* err, int32_t errlen);

#ifdef __cplusplus
}
#endif
#endif
