// Run-time specialisation of the LTE-controlled backward-Euler transient
// (kg_transient_lte) to one circuit topology.
//
// The generic kernel interprets the plan (gather lists, factor program,
// level lists) with one thread per small system, so every Newton iteration
// is a chain of dependent plan and parameter reads. Here the plan is written
// out as straight-line HIP: one lane integrates one instance, with its
// element values, matrix slots, solution, previous states and diode states
// in registers. The step controller, the Newton loop, the junction limiting
// and the stamps follow the generic kernel operation for operation (the same
// contributions summed in the same order, the same LU: multipliers
// l_ik = a_ik * (1 / u_kk) stored in place, the upper factor unscaled, the
// pivot reciprocal kept on the diagonal), so both produce the same accepted
// time grid up to the compiler's contractions.
#include <algorithm>
#include <cstdio>
#include <functional>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "kg_jit.h"

namespace kg {
namespace {

struct TV {
  int n = 0, nnz = 0, ne = 0;
  const int32_t *row_ptr, *col, *diag, *fprog_ptr, *fprog, *scon_ptr, *scon, *rcon_ptr, *rcon, *col_of, *isbr,
      *etype, *ep, *em, *ec;
};

TV tview(const PatLayout& L, const int32_t* b) {
  TV v;
  v.n = L.n;
  v.nnz = L.nnz;
  v.ne = L.ne;
  v.row_ptr = b + L.o_row_ptr;
  v.col = b + L.o_col;
  v.diag = b + L.o_diag;
  v.fprog_ptr = b + L.o_fprog_ptr;
  v.fprog = b + L.o_fprog;
  v.scon_ptr = b + L.o_scon_ptr;
  v.scon = b + L.o_scon;
  v.rcon_ptr = b + L.o_rcon_ptr;
  v.rcon = b + L.o_rcon;
  v.col_of = b + L.o_col_of;
  v.isbr = b + L.o_isbr;
  v.etype = b + L.o_etype;
  v.ep = b + L.o_ep;
  v.em = b + L.o_em;
  v.ec = b + L.o_ec;
  return v;
}

// Waveforms of the generic kernels (Pulse / Sinusoidal value_at,
// voltage_source.rs:50-93), the thermal voltage (constants.rs:1-4).
const char* kLtePrelude = R"(typedef long long i64;
#define KG_VT (1.380649e-23 * 300.0 / 1.602176634e-19)
__device__ __forceinline__ double wave(const double* p, double t) {
  const int kind = (int)p[2];
  const double* w = p + 3;
  if (kind == 1) {
    const double v1 = w[0], v2 = w[1], td = w[2], tr = w[3], tf = w[4], pw = w[5], per = w[6];
    if (t < td) return v1;
    const double tt = fmod(t - td, per);
    if (tt < tr) return v1 + (v2 - v1) * (tt / tr);
    if (tt < tr + pw) return v2;
    if (tt < tr + pw + tf) return v2 + (v1 - v2) * ((tt - tr - pw) / tf);
    return v1;
  }
  if (kind == 2) {
    const double vo = w[0], va = w[1], f = w[2], td = w[3], th = w[4], ph = w[5];
    if (t < td) return vo;
    const double om = 2.0 * 3.14159265358979323846 * f;
    const double phr = ph * (3.14159265358979323846 / 180.0);
    const double tt = t - td;
    return vo + va * exp(-th * tt) * sin(om * tt + phr);
  }
  return p[0];
}
)";

std::string S(int i) { return std::to_string(i); }

// KG_GUARD_STATIC as a literal of the generated code.
std::string fmt_static() {
  char b[32];
  std::snprintf(b, sizeof b, "%.17g", KG_GUARD_STATIC);
  return b;
}

// Value of permuted unknown i in vector `pre` (ground: 0.0), as the generic
// kernels write it: (p >= 0 ? x[p] : 0.0).
std::string node(const std::string& pre, int i) { return i >= 0 ? pre + S(i) : std::string("0.0"); }

// Value of a stamp contribution (the plan's gather codes) and of a
// right-hand-side contribution, as expressions over the kernel's names.
using CodeFn = std::string (*)(const TV& P, int e, int k);

// Emits, row by row in pivot order: the row's matrix slots v<s>
// (contributions in the plan's order, summed from zero as the generic
// kernels do), its right-hand side x<i>, its factorisation (the plan's factor
// program: multipliers l_ik = a_ik * (1 / u_kk) stored in place, the pivot
// reciprocal on the diagonal, `bad` set for a zero or non-finite pivot) and
// its forward substitution; then the back substitution. Each row only reads
// the upper factor of earlier rows, so a row's multipliers die with it and
// the arithmetic is that of factorising everything first (the same
// operations in the same order). before_row(i) may emit loads first;
// on_result(i) runs once x<i> is final (back substitution order).
//
// It also assigns the caller's `gb` the static-pivot guard's bound on the
// solve's normwise backward error (KG_GUARD_STATIC; the level-scheduled
// kernels' pivot_guard): guard_bound_coeff(m) * rho with
// rho = || |L||U| ||_inf / ||A||_inf. an_ accumulates ||A||_inf over the
// stamped rows, gr<i> is row i's |U| sum (its pivot and upper entries once
// the row is factored), gm_ the rows of |L||U|.
bool emit_solve(std::ostringstream& o, const TV& P, const char* ind, CodeFn slot, CodeFn exc, std::string& why,
                const std::function<void(int)>& before_row = nullptr,
                const std::function<void(int)>& on_result = nullptr) {
  const int n = P.n;
  int m = 0;
  for (int i = 0; i < n; i++) m = std::max(m, P.row_ptr[i + 1] - P.row_ptr[i]);
  char coeff[32];
  std::snprintf(coeff, sizeof coeff, "%.17g", guard_bound_coeff(m));
  o << ind << "double an_ = 0.0, gm_ = 0.0;\n";
  for (int i = 0; i < n; i++) {
    if (before_row) before_row(i);
    for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++) {
      o << ind << "double v" << s << " = 0.0";
      for (int q = P.scon_ptr[s]; q < P.scon_ptr[s + 1]; q++) {
        const std::string c = slot(P, P.scon[q] >> 3, P.scon[q] & 7);
        if (c.empty()) {
          why = "unknown stamp kind";
          return false;
        }
        o << " + (" << c << ")";
      }
      o << ";\n";
    }
    o << ind << "x" << i << " = 0.0";
    for (int q = P.rcon_ptr[i]; q < P.rcon_ptr[i + 1]; q++) {
      const std::string c = exc(P, P.rcon[q] >> 3, P.rcon[q] & 7);
      if (c.empty()) {
        why = "unknown excitation kind";
        return false;
      }
      o << " + (" << c << ")";
    }
    o << ";\n" << ind << "an_ = fmax(an_, 0.0";
    for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++) o << " + fabs(v" << s << ")";
    o << ");\n";
    const int32_t* fp = P.fprog + P.fprog_ptr[i];
    const int nk = *fp++;
    for (int kk = 0; kk < nk; kk++) {
      const int sik = fp[0], sdk = fp[1], nu = fp[2];
      fp += 3;
      if (sik < P.row_ptr[i] || sik >= P.diag[i]) {
        why = "factor program leaves the row";
        return false;
      }
      o << ind << "v" << sik << " = v" << sik << " * v" << sdk << ";\n";
      for (int u = 0; u < nu; u++, fp += 2) {
        if (fp[0] < P.row_ptr[i] || fp[0] >= P.row_ptr[i + 1]) {
          why = "factor program leaves the row";
          return false;
        }
        o << ind << "v" << fp[0] << " = v" << fp[0] << " - v" << sik << " * v" << fp[1] << ";\n";
      }
    }
    const int d = P.diag[i];
    o << ind << "const double gr" << i << " = fabs(v" << d << ")";
    for (int p = d + 1; p < P.row_ptr[i + 1]; p++) o << " + fabs(v" << p << ")";
    o << ";\n" << ind << "gm_ = fmax(gm_, gr" << i;
    for (int p = P.row_ptr[i]; p < d; p++) o << " + fabs(v" << p << ") * gr" << P.col[p];
    o << ");\n";
    // the running maxima are taken here, not sunk to their use at the end
    // (which would keep every row's pivot and multipliers live)
    o << ind << "asm volatile(\"\" :: \"v\"(an_), \"v\"(gm_));\n";
    o << ind << "bad |= !(fabs(v" << d << ") > 0.0) || !__builtin_isfinite(v" << d << ");\n"
      << ind << "v" << d << " = 1.0 / v" << d << ";\n";
    for (int p = P.row_ptr[i]; p < d; p++) o << ind << "x" << i << " = x" << i << " - v" << p << " * x" << P.col[p] << ";\n";
  }
  for (int i = n - 1; i >= 0; i--) {
    const int d = P.diag[i];
    for (int p = d + 1; p < P.row_ptr[i + 1]; p++) o << ind << "x" << i << " = x" << i << " - v" << p << " * x" << P.col[p] << ";\n";
    o << ind << "x" << i << " = x" << i << " * v" << d << ";\n";
    if (on_result) on_result(i);
  }
  o << ind << "gb = gm_ > 0.0 ? " << coeff << " * (gm_ / an_) : 0.0;\n";
  return true;
}

// Backward-Euler transient stamps (stampable.rs, the generic kernels' elem_y
// / elem_z / rhs_contrib at alpha = 1): names of the LTE kernel.
std::string lte_slot(const TV& P, int e, int k) {
  const int ty = P.etype[e];
  switch (k) {
    case CK_YP:
    case CK_YN: {
      std::string c = ty == KG_R ? "g" + S(e) : ty == KG_C ? "ch" + S(e) : ty == KG_D ? "gd" + S(e) : "0.0";
      return k == CK_YN ? "-" + c : c;
    }
    case CK_ONEP: return "1.0";
    case CK_ONEN: return "-1.0";
    case CK_BR: return ty == KG_L ? "-(lh" + S(e) + ")" : ty == KG_R ? "-r" + S(e) : "1.0";
    case CK_KP: return "g" + S(e);
    case CK_KN: return "-g" + S(e);
    default: return "";
  }
}
std::string lte_exc(const TV& P, int e, int k) {
  switch (k) {
    case RK_SRC: return P.etype[e] == KG_V ? "vs" + S(e) : "p" + S(e);
    case RK_DP: return "-id" + S(e);
    case RK_DM: return "id" + S(e);
    case RK_CP: return "-cq" + S(e);
    case RK_CM: return "cq" + S(e);
    case RK_IND: return "lq" + S(e);
    default: return "";
  }
}

// DC stamps (op.rs, dc.rs): capacitors are open, inductors shorts; a swept
// source takes the sweep value.
std::string dc_slot(const TV& P, int e, int k) {
  const int ty = P.etype[e];
  switch (k) {
    case CK_YP: return ty == KG_R ? "g" + S(e) : ty == KG_D ? "gd" + S(e) : "";
    case CK_YN: return ty == KG_R ? "-g" + S(e) : ty == KG_D ? "-gd" + S(e) : "";
    case CK_ONEP: return "1.0";
    case CK_ONEN: return "-1.0";
    case CK_BR: return ty == KG_R ? "-r" + S(e) : "";
    case CK_KP: return "g" + S(e);
    case CK_KN: return "-g" + S(e);
    default: return "";
  }
}
std::string dc_exc(const TV& P, int e, int k) {
  if (k == RK_DP) return "-id" + S(e);
  if (k == RK_DM) return "id" + S(e);
  if (k != RK_SRC) return "";
  return "(sweep_e == " + S(e) + " ? sv : p" + S(e) + ")";
}

}  // namespace

const char* jit_lte_name() { return "kg_lte_lane"; }

bool jit_lte_source(const PatLayout& L, const int32_t* b, std::string& src, std::string& why) {
  if (L.status != KG_OK) {
    why = "structurally singular pattern";
    return false;
  }
  TV P = tview(L, b);
  const int n = P.n, ne = P.ne;
  // registers held across a Newton iteration (2 per double): matrix slots,
  // right-hand side / solution, guess, two previous states, element values
  int nvals = 0, nd = 0;
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    nvals += ty == KG_V ? 10 : ty == KG_D ? 4 : 1;
    nd += ty == KG_D;
  }
  const int regs = 2 * (P.nnz + 5 * n + 3 * nd + nvals);
  if (regs > 400) {
    why = "per-lane state exceeds the register budget";
    return false;
  }
  std::ostringstream o;
  o << kLtePrelude;
  o << "extern \"C\" __global__ void __launch_bounds__(64) " << jit_lte_name()
    << "(const double* __restrict__ prm,\n"
       "    const double* __restrict__ ebase, const double* __restrict__ xop, const i64* __restrict__ xop_off,\n"
       "    double* __restrict__ tout, double* __restrict__ xout, const i64* __restrict__ xout_off,\n"
       "    int* __restrict__ npts, int* __restrict__ ntot, int* __restrict__ rejected, int* __restrict__ status,\n"
       "    int ninst, double tstop, double h0, double hmin, double hmax, double lte_rel, double lte_abs,\n"
       "    int max_points, double reltol, double iabstol, double vabstol, int maxiter) {\n"
       "  const int inst = (int)blockIdx.x * 64 + (int)threadIdx.x;\n"
       "  if (inst >= ninst) return;\n"
       "  int st = status[inst], cnt = 0, nt = ntot[inst], rej = 0;\n  double gb = 0.0;\n"
       "  if (st == 0) {\n";
  // element values, loaded once
  o << "  const double* PR = prm + (i64)inst * " << (int64_t)ne * KG_NPARAM << ";\n"
    << "  const double* EB = ebase + (i64)inst * " << ne << ";\n";
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    const std::string pe = "PR[" + S(e * KG_NPARAM);
    if (ty == KG_V) {
      o << "  const double pw" << e << "[10] = {";
      for (int k = 0; k < 10; k++) o << (k ? ", " : "") << "PR[" << e * KG_NPARAM + k << "]";
      o << "};\n";
    } else if (ty == KG_D) {
      // diode.rs:24-71 with SPICE3 junction limiting (the generic clean mode)
      o << "  const double is" << e << " = " << pe << "], nvt" << e << " = PR[" << e * KG_NPARAM + 1 << "] * KG_VT;\n"
        << "  const double vcrit" << e << " = nvt" << e << " * log(nvt" << e << " / (1.4142135623730951 * is" << e
        << "));\n";
    } else if (ty == KG_R || ty == KG_E || ty == KG_G || ty == KG_F || ty == KG_H) {
      o << "  const double g" << e << " = EB[" << e << "];\n";  // 1/R (value for G2 resistors), gains
      if (ty == KG_R) o << "  const double r" << e << " = " << pe << "];\n";
    } else {
      o << "  const double p" << e << " = " << pe << "];\n";  // C, L, I values
    }
  }
  // initial state: the operating point (permuted numbering), row 0 of the output
  o << "  const double* xin = xop + xop_off[inst];\n  double* o = xout + xout_off[inst];\n";
  for (int j = 0; j < n; j++) o << "  o[" << j << "] = xin[" << j << "];\n";
  for (int i = 0; i < n; i++) o << "  double xc" << i << " = xin[" << P.col_of[i] << "], xp" << i << " = xc" << i << ";\n";
  o << "  tout[(i64)inst * max_points] = 0.0;\n  cnt = 1;\n"
       "  double tcur = 0.0, h = h0, tprev = 0.0;\n  bool have_prev = false;\n"
       "  const i64 guard_max = 64 * (i64)max_points + 1000;\n"
       "  for (i64 guard = 0;; guard++) {\n"
    << "    if (guard > guard_max) { st = " << (int)KG_MAX_ITERATIONS << "; break; }\n"
    << "    h = fmin(h, hmax);\n    if (tcur + h > tstop) h = tstop - tcur;\n    const double tn = tcur + h;\n";
  // per-step values: sources at tn, companion conductances and currents
  // (stampable.rs:360-506, 676-817: i_eq = -(C/h) v_prev, -(L/h) i_prev)
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    if (ty == KG_V) o << "    const double vs" << e << " = wave(pw" << e << ", tn);\n";
    if (ty == KG_C) {
      o << "    const double ch" << e << " = 1.0 * (p" << e << " / h);\n"
        << "    const double cq" << e << " = -(ch" << e << ") * (" << node("xc", P.ep[e]) << " - " << node("xc", P.em[e])
        << ");\n";
    }
    if (ty == KG_L) {
      o << "    const double lh" << e << " = 1.0 * (p" << e << " / h);\n"
        << "    const double lq" << e << " = -(lh" << e << ") * xc" << P.ec[e] << ";\n";
    }
  }
  // Newton from the previous state
  for (int i = 0; i < n; i++) o << "    double g" << "u" << i << " = xc" << i << ", x" << i << " = 0.0;\n";
  for (int e = 0; e < ne; e++)
    if (P.etype[e] == KG_D)
      o << "    double vo" << e << " = " << node("xc", P.ep[e]) << " - " << node("xc", P.em[e]) << ", gd" << e
        << " = 0.0, id" << e << " = 0.0;\n";
  o << "    bool done = false;\n    int its = 0;\n"
       "    for (int it = 0; it < maxiter; it++) {\n      bool lim = false, bad = false;\n";
  for (int e = 0; e < ne; e++) {
    if (P.etype[e] != KG_D) continue;
    const std::string E = S(e);
    o << "      {\n        double v = " << node("gu", P.ep[e]) << " - " << node("gu", P.em[e]) << ";\n"
      << "        if (v > vcrit" << E << " && fabs(v - vo" << E << ") > nvt" << E << " + nvt" << E << ") {\n"
      << "          if (vo" << E << " > 0) {\n"
      << "            const double arg = 1.0 + (v - vo" << E << ") / nvt" << E << ";\n"
      << "            v = arg > 0 ? vo" << E << " + nvt" << E << " * log(arg) : vcrit" << E << ";\n"
      << "          } else {\n            v = nvt" << E << " * log(v / nvt" << E << ");\n          }\n"
      << "          lim = true;\n        }\n"
      << "        const double ex = exp(v / nvt" << E << ");\n"
      << "        gd" << E << " = is" << E << " / nvt" << E << " * ex;\n"
      << "        id" << E << " = is" << E << " * (ex - 1.0) - gd" << E << " * v;\n"
      << "        vo" << E << " = v;\n      }\n";
  }
  if (!emit_solve(o, P, "      ", lte_slot, lte_exc, why)) return false;
  // convergence (the clean-mode test of the generic kernels)
  o << "      bool cv = true;\n";
  for (int i = 0; i < n; i++)
    o << "      { const double df = fabs(x" << i << " - gu" << i << "), sc = fmax(fabs(x" << i << "), fabs(gu" << i
      << ")); if (!(df <= reltol * sc + " << (P.isbr[i] ? "iabstol" : "vabstol") << ")) cv = false; }\n";
  o << "      bool stop = " << (L.nonlinear ? "cv && !lim" : "true") << ";\n"
    << "      if (bad) { st = " << (int)KG_DECOMPOSITION << "; stop = true; }\n"
    << "      if (stop) { done = true; its = it + 1; break; }\n";
  for (int i = 0; i < n; i++) o << "      gu" << i << " = x" << i << ";\n";
  o << "    }\n"
    << "    if (done && !(gb <= " << fmt_static() << ")) st = " << (int)KG_DECOMPOSITION << ";  // static-pivot guard\n"
       "    if (st != 0) break;\n"
       "    if (!done) {  // Newton failure: retry with h / 8\n"
       "      rej++;\n      h /= 8.0;\n"
    << "      if (h < hmin) { st = " << (int)KG_MAX_ITERATIONS << "; break; }\n"
    << "      continue;\n    }\n"
       "    nt += its;\n"
       "    double ratio = 0.0;\n"
       "    if (have_prev) {  // local truncation error of capacitor voltages and inductor currents\n"
       "      const double hp = tcur - tprev;\n";
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    if (ty != KG_C && ty != KG_L) continue;
    o << "      {\n";
    if (ty == KG_C) {
      o << "        const double a = " << node("x", P.ep[e]) << " - " << node("x", P.em[e]) << ";\n"
        << "        const double b0 = " << node("xc", P.ep[e]) << " - " << node("xc", P.em[e]) << ";\n"
        << "        const double bm = " << node("xp", P.ep[e]) << " - " << node("xp", P.em[e]) << ";\n";
    } else {
      o << "        const double a = x" << P.ec[e] << ", b0 = xc" << P.ec[e] << ", bm = xp" << P.ec[e] << ";\n";
    }
    o << "        const double dd2 = 2.0 * ((a - b0) / h - (b0 - bm) / hp) / (h + hp);\n"
         "        const double lte = 0.5 * h * h * fabs(dd2);\n"
         "        const double tol = lte_rel * fmax(fabs(a), fabs(b0)) + lte_abs;\n"
         "        ratio = fmax(ratio, lte / tol);\n      }\n";
  }
  o << "    }\n"
       "    double hnext = h;\n"
       "    if (have_prev) {\n"
       "      if (ratio > 1.0) {\n"
       "        rej++;\n        h *= fmax(0.25, 0.9 / sqrt(ratio));\n"
    << "        if (h < hmin) { st = " << (int)KG_MAX_ITERATIONS << "; break; }\n"
    << "        continue;\n      }\n"
       "      hnext = ratio > 0.0 ? h * fmin(2.0, fmax(0.25, 0.9 / sqrt(ratio))) : 2.0 * h;\n"
       "    }\n"
    << "    if (cnt >= max_points) { st = " << (int)KG_UNEXPECTED << "; break; }\n";
  for (int i = 0; i < n; i++) o << "    xp" << i << " = xc" << i << ";\n    xc" << i << " = x" << i << ";\n";
  o << "    {\n      double* orow = o + (i64)cnt * " << n << ";\n";
  for (int i = 0; i < n; i++) o << "      orow[" << P.col_of[i] << "] = x" << i << ";\n";
  o << "    }\n"
       "    tout[(i64)inst * max_points + cnt] = tn;\n"
       "    cnt++;\n    tprev = tcur;\n    tcur = tn;\n    h = hnext;\n    have_prev = true;\n"
       "    if (!(tcur < tstop * (1.0 - 1e-12))) break;\n"
       "  }\n"
       "  }\n"
       "  status[inst] = st;\n  npts[inst] = cnt;\n  ntot[inst] = nt;\n  rejected[inst] = rej;\n}\n";
  src = o.str();
  return true;
}

}  // namespace kg

namespace kg {

const char* jit_dc_name() { return "kg_dc_lane"; }

// Linear operating points and DC sweeps (op.rs:15-86, dc.rs:14-124 for
// circuits without diodes: one solve per point, which krets' Newton loop
// accepts after its first pass): one lane per instance, its points in order.
bool jit_dc_newton_source(const PatLayout& L, const TV& P, std::string& src, std::string& why);

bool jit_dc_source(const PatLayout& L, const int32_t* b, std::string& src, std::string& why) {
  if (L.status != KG_OK) {
    why = "structurally singular pattern";
    return false;
  }
  if (L.nonlinear) return jit_dc_newton_source(L, tview(L, b), src, why);
  TV P = tview(L, b);
  const int n = P.n, ne = P.ne;
  // registers live across the solve: the upper factor and pivots, x, and the
  // element values loaded ahead of their rows
  int nu = 0;
  for (int i = 0; i < n; i++) nu += P.row_ptr[i + 1] - P.diag[i];
  if (2 * (nu + n) > 440) {
    why = "per-lane state exceeds the register budget";
    return false;
  }
  // element values each row reads, loaded `kAhead` rows before it (an empty
  // asm statement with a memory clobber keeps the compiler from hoisting
  // every load to the start)
  constexpr int kAhead = 8;
  std::vector<std::vector<std::pair<std::string, std::string>>> need(n);
  for (int i = 0; i < n; i++) {
    auto add = [&](const std::string& nm, const std::string& ld) { need[i].push_back({nm, ld}); };
    for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++)
      for (int q = P.scon_ptr[s]; q < P.scon_ptr[s + 1]; q++) {
        const int e = P.scon[q] >> 3, k = P.scon[q] & 7;
        if (k == CK_YP || k == CK_YN || k == CK_KP || k == CK_KN) add("g" + S(e), "EB[" + S(e) + "]");
        if (k == CK_BR && P.etype[e] == KG_R) add("r" + S(e), "PR[" + S(e * KG_NPARAM) + "]");
      }
    for (int q = P.rcon_ptr[i]; q < P.rcon_ptr[i + 1]; q++) {
      const int e = P.rcon[q] >> 3;
      add("p" + S(e), "PR[" + S(e * KG_NPARAM) + "]");
    }
  }
  std::ostringstream o;
  std::set<std::string> have;
  auto issue = [&](int r) {
    if (r >= n) return;
    bool any = false;
    for (auto& v : need[r])
      if (have.insert(v.first).second) {
        o << "    const double " << v.first << " = " << v.second << ";\n";
        any = true;
      }
    if (any) o << "    asm volatile(\"\" ::: \"memory\");\n";
  };
  o << "typedef long long i64;\n"
    << "extern \"C\" __global__ void __launch_bounds__(64) " << jit_dc_name()
    << "(const double* __restrict__ prm,\n"
       "    const double* __restrict__ ebase, int ninst, int sweep_e, double start, double step, int npts,\n"
       "    double* __restrict__ out, const i64* __restrict__ out_off, int* __restrict__ status,\n"
       "    int* __restrict__ iters, int compat, double reltol, double iabstol, double vabstol, int maxiter) {\n"
       "  const int inst = (int)blockIdx.x * 64 + (int)threadIdx.x;\n"
       "  if (inst >= ninst) return;\n"
    << "  const double* PR = prm + (i64)inst * " << (int64_t)ne * KG_NPARAM << ";\n"
    << "  const double* EB = ebase + (i64)inst * " << ne << ";\n"
    << "  int st = 0;\n  double gb = 0.0;\n";
  for (int i = 0; i < n; i++) o << "  double x" << i << ";\n";
  o << "  for (int pt = 0; pt < npts && st == 0; pt++) {\n"
       "    const double sv = start + (double)pt * step;\n"
       "    bool bad = false;\n";
  for (int r = 0; r < kAhead; r++) issue(r);
  // each unknown is written as soon as the back substitution has it (no solution vector held to the end)
  o << "    double* o = out + out_off[inst] + (i64)pt * " << n << ";\n";
  if (!emit_solve(o, P, "    ", dc_slot, dc_exc, why, [&](int i) { issue(i + kAhead); },
                  [&](int i) { o << "    o[" << P.col_of[i] << "] = x" << i << ";\n"; }))
    return false;
  o << "    if (bad || !(gb <= " << fmt_static() << ")) st = " << (int)KG_DECOMPOSITION << ";\n";
  o << "    if (iters) iters[(i64)inst * npts + pt] = 1;\n"
       "  }\n"
       "  status[inst] = st;\n}\n";
  src = o.str();
  return true;
}

}  // namespace kg

namespace kg {

const char* jit_tran_name() { return "kg_tran_lane"; }

// Fixed-step transient (transient.rs:14-114; the generic k_tran): backward
// Euler with krets' quirks (method 0: sources at the previous step's time on
// the first Newton iteration, krets' diode companion with its 0.5 V default
// for an absent plus node, the bookkeeping time key in the first
// convergence test), clean backward Euler (1) and trapezoidal (2) with
// SPICE3 junction limiting. One lane integrates one instance; the Newton
// loop, companions and stamps are those of k_tran, operation for operation.
bool jit_tran_source(const PatLayout& L, const int32_t* b, std::string& src, std::string& why) {
  if (L.status != KG_OK) {
    why = "structurally singular pattern";
    return false;
  }
  TV P = tview(L, b);
  const int n = P.n, ne = P.ne;
  int nvals = 0, nd = 0, nu = 0;
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    nvals += ty == KG_V ? 12 : ty == KG_D ? 5 : ty == KG_C || ty == KG_L ? 4 : 1;
    nd += ty == KG_D;
  }
  for (int i = 0; i < n; i++) nu += P.row_ptr[i + 1] - P.diag[i];
  if (2 * (nu + 4 * n + 3 * nd + nvals) > 400) {
    why = "per-lane state exceeds the register budget";
    return false;
  }
  std::ostringstream o;
  o << kLtePrelude;
  o << "extern \"C\" __global__ void __launch_bounds__(64) " << jit_tran_name()
    << "(const double* __restrict__ prm,\n"
       "    const double* __restrict__ ebase, const double* __restrict__ xop, const i64* __restrict__ xop_off,\n"
       "    double* __restrict__ out, const i64* __restrict__ out_off, int* __restrict__ status,\n"
       "    int* __restrict__ iters, double* __restrict__ err, int ninst, double h, double alpha, int nsteps,\n"
       "    int method, double reltol, double iabstol, double vabstol, int maxiter) {\n"
       "  const int inst = (int)blockIdx.x * 64 + (int)threadIdx.x;\n"
       "  if (inst >= ninst) return;\n"
       "  int st = status[inst];\n  double gb = 0.0, gbmax = 0.0;\n"
       "  const bool compat = method == 0, trap = method == 2;\n"
       "  if (st == 0) {\n"
    << "  const double* PR = prm + (i64)inst * " << (int64_t)ne * KG_NPARAM << ";\n"
    << "  const double* EB = ebase + (i64)inst * " << ne << ";\n";
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    const std::string pe = "PR[" + S(e * KG_NPARAM);
    if (ty == KG_V) {
      o << "  const double pw" << e << "[10] = {";
      for (int k = 0; k < 10; k++) o << (k ? ", " : "") << "PR[" << e * KG_NPARAM + k << "]";
      o << "};\n";
    } else if (ty == KG_D) {
      // diode.rs:24-71: krets' limit (effectively none) and the clean SPICE3 limit
      o << "  const double is" << e << " = " << pe << "], nd" << e << " = PR[" << e * KG_NPARAM + 1 << "];\n"
        << "  const double nvt" << e << " = nd" << e << " * KG_VT;\n"
        << "  const double vcrit" << e << " = nvt" << e << " * log(nvt" << e << " / (1.4142135623730951 * is" << e
        << "));\n"
        << "  const double vcritc" << e << " = nvt" << e << " * log(1.7976931348623157e308 * nd" << e
        << " * KG_VT / is" << e << ");\n";
    } else if (ty == KG_R || ty == KG_E || ty == KG_G || ty == KG_F || ty == KG_H) {
      o << "  const double g" << e << " = EB[" << e << "];\n";
      if (ty == KG_R) o << "  const double r" << e << " = " << pe << "];\n";
    } else {
      o << "  const double p" << e << " = " << pe << "];\n";
    }
  }
  o << "  const double* xin = xop + xop_off[inst];\n  double* o = out + out_off[inst];\n";
  for (int j = 0; j < n; j++) o << "  o[" << j << "] = xin[" << j << "];\n";
  for (int i = 0; i < n; i++) o << "  double xp" << i << " = xin[" << P.col_of[i] << "], x" << i << " = 0.0;\n";
  o << "  if (iters) iters[(i64)inst * (nsteps + 1)] = 0;\n";
  // trapezoidal companion states: capacitor currents 0, inductor voltages v(x0)
  for (int e = 0; e < ne; e++) {
    if (P.etype[e] == KG_C) o << "  double edc" << e << " = 0.0;\n";
    if (P.etype[e] == KG_L) o << "  double edl" << e << " = " << node("xp", P.ep[e]) << " - " << node("xp", P.em[e]) << ";\n";
  }
  o << "  for (int s = 1; s <= nsteps; s++) {\n"
       "    const double tn = (double)s * h, tprev = (double)(s - 1) * h;\n";
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    if (ty == KG_V) o << "    const double vsn" << e << " = wave(pw" << e << ", tn), vsp" << e << " = wave(pw" << e << ", tprev);\n";
    if (ty == KG_C)
      o << "    const double ch" << e << " = alpha * (p" << e << " / h);\n"
        << "    double cq" << e << " = -(ch" << e << ") * (" << node("xp", P.ep[e]) << " - " << node("xp", P.em[e]) << ");\n"
        << "    if (alpha != 1.0) cq" << e << " -= edc" << e << ";\n";
    if (ty == KG_L)
      o << "    const double lh" << e << " = alpha * (p" << e << " / h);\n"
        << "    double lq" << e << " = -(lh" << e << ") * xp" << P.ec[e] << ";\n"
        << "    if (alpha != 1.0) lq" << e << " -= edl" << e << ";\n";
  }
  for (int i = 0; i < n; i++) o << "    double gu" << i << " = xp" << i << ";\n";
  for (int e = 0; e < ne; e++)
    if (P.etype[e] == KG_D)
      o << "    double vo" << e << " = " << node("xp", P.ep[e]) << " - " << node("xp", P.em[e]) << ", gd" << e
        << " = 0.0, id" << e << " = 0.0;\n";
  o << "    bool done = false;\n    int used = 1;\n"
       "    for (int it = 0; it < maxiter; it++) {\n      bool lim = false, bad = false;\n";
  for (int e = 0; e < ne; e++)
    if (P.etype[e] == KG_V) o << "      const double vs" << e << " = (compat && it == 0) ? vsp" << e << " : vsn" << e << ";\n";
  for (int e = 0; e < ne; e++) {
    if (P.etype[e] != KG_D) continue;
    const std::string E = S(e);
    const int p = P.ep[e], m = P.em[e];
    o << "      if (compat) {  // an absent plus node reads 0.5 V, an absent minus node 0 V (diode.rs:26-36)\n"
      << "        const double vd = " << (p >= 0 ? "gu" + S(p) : std::string("0.5")) << " - "
      << (m >= 0 ? "gu" + S(m) : std::string("0.0")) << ";\n"
      << "        const double vl = fmin(fmax(vd, -vcritc" << E << "), vcritc" << E << ");\n"
      << "        const double ex = exp(vl / nvt" << E << ");\n"
      << "        gd" << E << " = (is" << E << " / nvt" << E << ") * ex;\n"
      << "        id" << E << " = is" << E << " * (ex - 1.0) - gd" << E << " * vd;\n"
      << "      } else {\n"
      << "        double v = " << node("gu", p) << " - " << node("gu", m) << ";\n"
      << "        if (v > vcrit" << E << " && fabs(v - vo" << E << ") > nvt" << E << " + nvt" << E << ") {\n"
      << "          if (vo" << E << " > 0) {\n"
      << "            const double arg = 1.0 + (v - vo" << E << ") / nvt" << E << ";\n"
      << "            v = arg > 0 ? vo" << E << " + nvt" << E << " * log(arg) : vcrit" << E << ";\n"
      << "          } else {\n            v = nvt" << E << " * log(v / nvt" << E << ");\n          }\n"
      << "          lim = true;\n        }\n"
      << "        const double ex = exp(v / nvt" << E << ");\n"
      << "        gd" << E << " = is" << E << " / nvt" << E << " * ex;\n"
      << "        id" << E << " = is" << E << " * (ex - 1.0) - gd" << E << " * v;\n"
      << "        vo" << E << " = v;\n      }\n";
  }
  if (!emit_solve(o, P, "      ", lte_slot, lte_exc, why)) return false;
  o << "      if (bad) { st = " << (int)KG_DECOMPOSITION << "; break; }\n";
  if (L.nonlinear) {
    o << "      bool cv = true;\n";
    for (int i = 0; i < n; i++)
      o << "      { const double df = fabs(x" << i << " - gu" << i << "), sc = fmax(fabs(x" << i << "), fabs(gu" << i
        << ")); if (!(df <= reltol * sc + " << (P.isbr[i] ? "iabstol" : "vabstol") << ")) cv = false; }\n";
    o << "      if (compat && it == 0)  // transient.rs:204, the time key t_n against t_{n-1}\n"
         "        cv = cv && fabs(tn - tprev) <= reltol * fmax(fabs(tn), fabs(tprev)) + vabstol;\n"
         "      if (!compat && lim) cv = false;\n"
         "      if (cv) { done = true; used = it + 1; break; }\n";
    for (int i = 0; i < n; i++) o << "      gu" << i << " = x" << i << ";\n";
  } else {
    o << "      done = true;\n      break;\n";
  }
  o << "    }\n"
    << "    if (done) {  // static-pivot guard\n      gbmax = fmax(gbmax, gb);\n"
    << "      if (!(gb <= " << fmt_static() << ")) st = " << (int)KG_DECOMPOSITION << ";\n    }\n"
       "    if (st != 0) break;\n"
    << "    if (!done) { st = " << (int)KG_MAX_ITERATIONS << "; break; }\n"
    << "    if (trap) {\n";
  for (int e = 0; e < ne; e++) {
    if (P.etype[e] == KG_C)
      o << "      edc" << e << " = 2.0 * p" << e << " / h * ((" << node("x", P.ep[e]) << " - " << node("x", P.em[e])
        << ") - (" << node("xp", P.ep[e]) << " - " << node("xp", P.em[e]) << ")) - edc" << e << ";\n";
    if (P.etype[e] == KG_L) o << "      edl" << e << " = " << node("x", P.ep[e]) << " - " << node("x", P.em[e]) << ";\n";
  }
  o << "    }\n    double* orow = o + (i64)s * " << n << ";\n";
  for (int i = 0; i < n; i++) o << "    orow[" << P.col_of[i] << "] = x" << i << ";\n    xp" << i << " = x" << i << ";\n";
  o << "    if (iters) iters[(i64)inst * (nsteps + 1) + s] = used;\n"
       "  }\n"
       "  }\n"
       "  status[inst] = st;\n"
       "  if (err) err[inst] = gbmax;\n}\n";
  src = o.str();
  return true;
}

}  // namespace kg

namespace kg {

// Operating points and DC sweeps of circuits with diodes (op.rs:15-86,
// dc.rs:14-124; the generic k_newton_dc): one lane per instance walks its
// points, each a Newton loop warm-started from the previous point. compat
// (krets): the first iteration of the first point sees an empty solution map
// (diodes at 0.5 V / 0 V, no convergence), and the first iteration of every
// later sweep point compares the bookkeeping step key; clean: zero start and
// SPICE3 junction limiting.
bool jit_dc_newton_source(const PatLayout& L, const TV& P, std::string& src, std::string& why) {
  const int n = P.n, ne = P.ne;
  int nvals = 0, nd = 0, nu = 0;
  for (int e = 0; e < ne; e++) {
    nvals += P.etype[e] == KG_D ? 5 : 2;
    nd += P.etype[e] == KG_D;
  }
  for (int i = 0; i < n; i++) nu += P.row_ptr[i + 1] - P.diag[i];
  if (2 * (nu + 3 * n + 3 * nd + nvals) > 400) {
    why = "per-lane state exceeds the register budget";
    return false;
  }
  std::ostringstream o;
  o << "typedef long long i64;\n#define KG_VT (1.380649e-23 * 300.0 / 1.602176634e-19)\n"
    << "extern \"C\" __global__ void __launch_bounds__(64) " << jit_dc_name()
    << "(const double* __restrict__ prm,\n"
       "    const double* __restrict__ ebase, int ninst, int sweep_e, double start, double step, int npts,\n"
       "    double* __restrict__ out, const i64* __restrict__ out_off, int* __restrict__ status,\n"
       "    int* __restrict__ iters, int compat, double reltol, double iabstol, double vabstol, int maxiter) {\n"
       "  const int inst = (int)blockIdx.x * 64 + (int)threadIdx.x;\n"
       "  if (inst >= ninst) return;\n"
    << "  const double* PR = prm + (i64)inst * " << (int64_t)ne * KG_NPARAM << ";\n"
    << "  const double* EB = ebase + (i64)inst * " << ne << ";\n";
  for (int e = 0; e < ne; e++) {
    const int ty = P.etype[e];
    if (ty == KG_V || ty == KG_I) o << "  const double p" << e << " = PR[" << e * KG_NPARAM << "];\n";
    if (ty == KG_R || ty == KG_E || ty == KG_G || ty == KG_F || ty == KG_H)
      o << "  const double g" << e << " = EB[" << e << "];\n";
    if (ty == KG_R) o << "  const double r" << e << " = PR[" << e * KG_NPARAM << "];\n";
    if (ty == KG_D)
      o << "  const double is" << e << " = PR[" << e * KG_NPARAM << "], nd" << e << " = PR[" << e * KG_NPARAM + 1
        << "];\n"
        << "  const double nvt" << e << " = nd" << e << " * KG_VT;\n"
        << "  const double vcrit" << e << " = nvt" << e << " * log(nvt" << e << " / (1.4142135623730951 * is" << e
        << "));\n"
        << "  const double vcritc" << e << " = nvt" << e << " * log(1.7976931348623157e308 * nd" << e
        << " * KG_VT / is" << e << ");\n";
  }
  o << "  int st = 0;\n  double gb = 0.0;\n";
  for (int i = 0; i < n; i++) o << "  double gu" << i << " = 0.0, x" << i << " = 0.0;\n";
  o << "  for (int pt = 0; pt < npts && st == 0; pt++) {\n"
       "    const double sv = start + (double)pt * step;\n"
       "    bool prev_empty = compat && pt == 0;\n";
  for (int e = 0; e < ne; e++)
    if (P.etype[e] == KG_D)
      o << "    double vo" << e << " = compat ? 0.0 : " << node("gu", P.ep[e]) << " - " << node("gu", P.em[e]) << ", gd"
        << e << " = 0.0, id" << e << " = 0.0;\n";
  o << "    bool done = false;\n    int used = maxiter;\n"
       "    for (int it = 0; it < maxiter; it++) {\n      bool lim = false, bad = false;\n";
  for (int e = 0; e < ne; e++) {
    if (P.etype[e] != KG_D) continue;
    const std::string E = S(e);
    const int p = P.ep[e], m = P.em[e];
    o << "      if (compat) {  // an absent key reads 0.5 V (plus) / 0 V (minus) (diode.rs:26-36)\n"
      << "        const double vd = (prev_empty" << (p < 0 ? " || true" : "") << " ? 0.5 : "
      << (p >= 0 ? "gu" + S(p) : std::string("0.5")) << ") - (prev_empty" << (m < 0 ? " || true" : "")
      << " ? 0.0 : " << (m >= 0 ? "gu" + S(m) : std::string("0.0")) << ");\n"
      << "        const double vl = fmin(fmax(vd, -vcritc" << E << "), vcritc" << E << ");\n"
      << "        const double ex = exp(vl / nvt" << E << ");\n"
      << "        gd" << E << " = (is" << E << " / nvt" << E << ") * ex;\n"
      << "        id" << E << " = is" << E << " * (ex - 1.0) - gd" << E << " * vd;\n"
      << "      } else {\n"
      << "        double v = " << node("gu", p) << " - " << node("gu", m) << ";\n"
      << "        if (v > vcrit" << E << " && fabs(v - vo" << E << ") > nvt" << E << " + nvt" << E << ") {\n"
      << "          if (vo" << E << " > 0) {\n"
      << "            const double arg = 1.0 + (v - vo" << E << ") / nvt" << E << ";\n"
      << "            v = arg > 0 ? vo" << E << " + nvt" << E << " * log(arg) : vcrit" << E << ";\n"
      << "          } else {\n            v = nvt" << E << " * log(v / nvt" << E << ");\n          }\n"
      << "          lim = true;\n        }\n"
      << "        const double ex = exp(v / nvt" << E << ");\n"
      << "        gd" << E << " = is" << E << " / nvt" << E << " * ex;\n"
      << "        id" << E << " = is" << E << " * (ex - 1.0) - gd" << E << " * v;\n"
      << "        vo" << E << " = v;\n      }\n";
  }
  if (!emit_solve(o, P, "      ", dc_slot, dc_exc, why)) return false;
  o << "      bool stop = bad;\n"
       "      if (!prev_empty) {\n        bool cv = true;\n";
  for (int i = 0; i < n; i++)
    o << "        { const double df = fabs(x" << i << " - gu" << i << "), sc = fmax(fabs(x" << i << "), fabs(gu" << i
      << ")); if (!(df <= reltol * sc + " << (P.isbr[i] ? "iabstol" : "vabstol") << ")) cv = false; }\n";
  o << "        if (compat && it == 0 && pt > 0 && sweep_e >= 0) {  // the step key: pt against pt - 1\n"
       "          const double v1 = (double)pt, v0 = (double)(pt - 1);\n"
       "          cv = cv && fabs(v1 - v0) <= reltol * fmax(fabs(v1), fabs(v0)) + vabstol;\n"
       "        }\n"
       "        if (!compat && lim) cv = false;\n"
       "        stop = stop || cv;\n"
       "      }\n"
       "      prev_empty = false;\n"
    << "      if (bad) st = " << (int)KG_DECOMPOSITION << ";\n"
    << "      if (stop) { done = true; used = it + 1; break; }\n";
  for (int i = 0; i < n; i++) o << "      gu" << i << " = x" << i << ";\n";
  o << "    }\n"
    << "    if (!done) st = " << (int)KG_MAX_ITERATIONS << ";\n"
    << "    else if (st == 0 && !(gb <= " << fmt_static() << ")) st = " << (int)KG_DECOMPOSITION << ";  // static-pivot guard\n"
    << "    double* o = out + out_off[inst] + (i64)pt * " << n << ";\n";
  for (int i = 0; i < n; i++) o << "    o[" << P.col_of[i] << "] = x" << i << ";\n    gu" << i << " = x" << i << ";\n";
  o << "    if (iters) iters[(i64)inst * npts + pt] = used;\n"
       "  }\n"
       "  status[inst] = st;\n}\n";
  src = o.str();
  return true;
}

}  // namespace kg
