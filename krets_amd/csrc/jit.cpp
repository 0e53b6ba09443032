// Run-time specialisation of the AC sweep to one circuit topology.
//
// The host symbolic pass fixes, for a shared topology, every LU slot, the
// static pivot order and which element contributions land in each slot.
// Instead of interpreting that plan on the device (the level-scheduled
// kernels read it from memory), this file writes it out as straight-line
// HIP: one lane solves one (instance, frequency) system with the matrix, the
// factors and the solution in registers.
//
// Factorisation (row i in pivot order, up-looking as in the plan's factor
// program) keeps the upper factor scaled by the reciprocal pivot, U' =
// diag(U)^-1 U, and the forward-substitution vector scaled the same way, so
// that neither the lower factor nor the pivots outlive their row:
//   a_ij -= a_ik * U'_kj              (k < i in the row's pattern, j > k)
//   y_i   = b_i - sum_k a_ik * y'_k
//   U'_ij = a_ij / a_ii,  y'_i = y_i / a_ii
//   x_i   = y'_i - sum_j U'_ij x_j     (back substitution, i descending)
// This is the same LU (the same pivots, the same fill) as the plan's factor
// program, with the multiplier division moved onto the upper factor.
//
// Mixed precision repeats the factorisation in complex64 on every pass (it
// costs the same VALU work as re-reading stored factors would cost in
// registers this kernel does not have), with FP64 residuals b - A x from the
// FP64 stamps and FP64 accumulation of the corrections, as the generic
// kernels do.
#include "kg_jit.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <utility>
#include <vector>

namespace kg {
namespace {

struct View {
  int n = 0, nnz = 0, ne = 0;
  const int32_t *row_ptr, *col, *diag, *fprog_ptr, *fprog, *scon_ptr, *scon, *rcon_ptr, *rcon, *col_of, *etype;
};

View view(const PatLayout& L, const int32_t* b) {
  View v;
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
  v.etype = b + L.o_etype;
  return v;
}

const char* kPrelude = R"(typedef long long i64;
struct __attribute__((aligned(16))) Z { double re, im; };
struct __attribute__((aligned(8))) W { float re, im; };
__device__ __forceinline__ Z operator+(Z a, Z b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Z operator-(Z a, Z b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Z operator*(Z a, Z b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ W operator-(W a, W b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ W operator*(W a, W b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Z cz(Z a) { return a; }
__device__ __forceinline__ Z cz(W a) { return {(double)a.re, (double)a.im}; }
__device__ __forceinline__ W cw(Z a) { return {(float)a.re, (float)a.im}; }
__device__ __forceinline__ Z scal(double s, Z a) { return {s * a.re, s * a.im}; }
// product with a factor known to be real (a stamp with no imaginary part that
// no elimination step has changed): two multiplications where the complex
// product spends four plus two additions (IEEE arithmetic keeps its
// 0 * x.im terms)
__device__ __forceinline__ Z rmul(double a, Z x) { return {a * x.re, a * x.im}; }
__device__ __forceinline__ W rmul(float a, W x) { return {a * x.re, a * x.im}; }
__device__ __forceinline__ double m2(Z a) { return a.re * a.re + a.im * a.im; }
// running max norm over components, max(|re|, |im|), of complex64 values: one
// instruction per value, and no squares that could leave the FP32 range
__device__ __forceinline__ float cmax(float m, W a) {
  return __builtin_fmaxf(m, __builtin_fmaxf(__builtin_fabsf(a.re), __builtin_fabsf(a.im)));
}
__device__ __forceinline__ bool fin(Z a) { return __builtin_isfinite(a.re) && __builtin_isfinite(a.im); }
// |re| + |im|: the entry norm of the growth guard (submultiplicative, within
// a factor sqrt(2) of the modulus)
__device__ __forceinline__ double c1(Z a) { return __builtin_fabs(a.re) + __builtin_fabs(a.im); }
__device__ __forceinline__ double c1(W a) { return (double)__builtin_fabsf(a.re) + (double)__builtin_fabsf(a.im); }
// queues system `sys` for the dense partial-pivoting fixup (queue words: [0]
// count, [1 + q] systems); a full queue marks the instance instead
__device__ __forceinline__ void fix_push(unsigned long long* fixq, i64 cap, i64 sys, int* status, int inst) {
  const unsigned long long q = atomicAdd(fixq, 1ull);
  if (q < (unsigned long long)cap) fixq[1 + q] = (unsigned long long)sys;
  else atomicMax(&status[inst], 2);
}
// reciprocal of a pivot, conj(b) / |b|^2, without branches: an FP64 pivot is
// first scaled by the power of two of its larger component (exact), so
// |b|^2 cannot leave the FP64 range; an FP32 pivot's |b|^2 always fits FP64
__device__ __forceinline__ Z rcpF(Z b) {
  const int k = __builtin_amdgcn_frexp_exp(__builtin_fmax(__builtin_fabs(b.re), __builtin_fabs(b.im)));
  const double re = __builtin_amdgcn_ldexp(b.re, -k), im = __builtin_amdgcn_ldexp(b.im, -k);
  // m in [0.25, 2): hardware reciprocal estimate and two Newton steps (a
  // shorter dependent chain than a correctly rounded division; ~1 ulp)
  const double m = re * re + im * im;
  double q = __builtin_amdgcn_rcp(m);
  q = q * (2.0 - m * q);
  q = q * (2.0 - m * q);
  return {__builtin_amdgcn_ldexp(re * q, -k), __builtin_amdgcn_ldexp(-im * q, -k)};
}
// complex64 pivots: |b|^2 in FP64 (cannot leave its range) and the hardware
// reciprocal estimate as it is: its relative error (measured on gfx950: at
// most 4.6e-8, 2^-24.4) is that of the FP32 rounding that follows, and the
// refinement corrects the complex64 factorisation anyway (a Newton step
// costs two more dependent instructions per pivot and 1.5% of the kernel)
__device__ __forceinline__ W rcpF(W b) {
  const double re = b.re, im = b.im, m = re * re + im * im;
  const double q = __builtin_amdgcn_rcp(m);
  return {(float)(re * q), (float)(-im * q)};
}
// block order that keeps consecutive systems (one instance) on one XCD
__device__ __forceinline__ i64 xcd_block() {
  const i64 nb = gridDim.x, b = blockIdx.x;
  const i64 q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}
)";

// Emits straight-line code. Element values are loaded once per pass (a pass
// is a brace scope), `lookahead` rows before the row that first uses them, and
// an empty asm statement with a memory clobber after each row's loads keeps
// the compiler from hoisting every load of the kernel to its start (which
// would hold all element values in registers at once).

// refinement passes load x as well as element values: a shorter distance
// keeps the values in flight within the register budget
#ifndef KG_LANE_AHEAD_REFINE
#define KG_LANE_AHEAD_REFINE 4
#endif
#ifndef KG_LANE_AHEAD
#define KG_LANE_AHEAD 8
#endif
// rows between scheduling barriers in the forward (factor and forward
// substitution) and backward sweeps; 0: none
#ifndef KG_LANE_FWD_BARRIER
#define KG_LANE_FWD_BARRIER 1
#endif
#ifndef KG_LANE_BWD_BARRIER
#define KG_LANE_BWD_BARRIER 4
#endif
int LOOKAHEAD_REFINE = KG_LANE_AHEAD_REFINE;

// A double as a C++ literal that reads back to the same value.
std::string fmt_double(double v) {
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

struct Gen {
  const View& P;
  std::ostringstream out;         // kernel text
  std::ostringstream* cur = &out; // where statements go
  std::vector<std::string> loads; // loads first needed by the row being generated
  std::set<int> have_g, have_p, have_x;
  std::string why;
  bool track_bm = false;          // mixed precision: |b|_inf^2 for the correction scale
  // mixed precision with the solution split as x = x0 + dx: x0 (the complex64
  // first solution) in LDS, dx (complex64 sum of the corrections) in registers
  bool xsplit = false;
  // element values from the table blocked by 64 instances (entry
  // (inst / 64 * ne + e) * 64 + inst % 64), so that a wave of consecutive
  // instances reads one contiguous run per value
  bool emajor = false;
  int lookahead = KG_LANE_AHEAD;  // rows between a load and the row that first uses it
  // names of element-value loads
  std::string gname = "g", pname = "p";
  // growth guard (first pass): an = ||A||_inf and gm = || |L||U| ||_inf in
  // |re| + |im| entry norms, the latter bounded from above through
  // |l_ik| <= |v_ik| |1/u_kk| (c1 is submultiplicative)
  bool growth = false;
  std::function<void(std::ostringstream&, int)> on_result;
  explicit Gen(const View& v) : P(v) {}

  // Element values come from the per-instance value table (1/R for plain
  // resistors, the element value otherwise: C, L, G2 resistors, current
  // sources), whose small offsets fit the loads' immediate field; only the
  // AC magnitude of voltage sources is read from the parameter records.
  std::string g(int e) {
    std::string nm = gname + std::to_string(e);
    if (have_g.insert(e).second) {
      if (emajor)  // buffer load: the wave's block (scalar), the lane's offset, a constant per element
        loads.push_back("    const double " + nm + " = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(EBR, jl8, " +
                        std::to_string(512 * e) + ", 0));\n");
      else
        loads.push_back("    const double " + nm + " = EB[" + std::to_string(e) + "];\n");
    }
    return nm;
  }
  std::string p(int e, int k) {  // params[e][k]
    std::string nm = pname + std::to_string(e) + "_" + std::to_string(k);
    if (have_p.insert(e * KG_NPARAM + k).second)
      loads.push_back("    const double " + nm + " = PR[" + std::to_string(e * KG_NPARAM + k) + "];\n");
    return nm;
  }
  // FP64 value of LU slot s as (re, im) expressions; both empty when the slot
  // receives no contribution. Follows the stamp kinds of the plan's gather
  // lists (AC stamps, stampable.rs): +-y (1/R, j w C), +-1, branch diagonal,
  // +-k (controlled-source gains).
  bool slot(int s, std::string& re, std::string& im) {
    std::vector<std::string> R, I;
    for (int q = P.scon_ptr[s]; q < P.scon_ptr[s + 1]; q++) {
      int code = P.scon[q], e = code >> 3, k = code & 7, ty = P.etype[e];
      switch (k) {
        case CK_YP:
        case CK_YN: {
          const char* sg = k == CK_YN ? "-" : "";
          if (ty == KG_R) R.push_back(sg + g(e));
          else if (ty == KG_C) I.push_back(std::string(sg) + "(w * " + g(e) + ")");
          else if (ty == KG_D) { why = "diode small-signal stamps are not generated"; return false; }
          break;
        }
        case CK_ONEP: R.push_back("1.0"); break;
        case CK_ONEN: R.push_back("-1.0"); break;
        case CK_BR:
          if (ty == KG_L) I.push_back("-(w * " + g(e) + ")");
          else if (ty == KG_R) R.push_back("-" + g(e));
          else R.push_back("1.0");
          break;
        case CK_KP: R.push_back(g(e)); break;  // controlled-source gain
        case CK_KN: R.push_back("-" + g(e)); break;
        default: why = "unknown stamp kind"; return false;
      }
    }
    auto join = [](const std::vector<std::string>& v) {
      std::string s;
      for (size_t i = 0; i < v.size(); i++) s += (i ? " + " : "") + v[i];
      return s;
    };
    re = join(R);
    im = join(I);
    return true;
  }
  bool has_slot(int s) const { return P.scon_ptr[s + 1] > P.scon_ptr[s]; }
  // FP64 right-hand side of row i (AC excitation, ac.rs: V sources carry their
  // AC magnitude, current sources their DC value); empty when zero.
  bool rhs(int i, std::string& re) {
    std::vector<std::string> R;
    for (int q = P.rcon_ptr[i]; q < P.rcon_ptr[i + 1]; q++) {
      int code = P.rcon[q], e = code >> 3, k = code & 7, ty = P.etype[e];
      if (k != RK_SRC) { why = "only source excitations are generated"; return false; }
      R.push_back(ty == KG_V ? p(e, 1) : g(e));
    }
    re.clear();
    for (size_t t = 0; t < R.size(); t++) re += (t ? " + " : "") + R[t];
    return true;
  }
  static std::string zexpr(const std::string& re, const std::string& im) {
    return "Z{" + (re.empty() ? std::string("0.0") : re) + ", " + (im.empty() ? std::string("0.0") : im) + "}";
  }
  static std::string conv(const std::string& F, const std::string& zval) {
    return F == "Z" ? zval : "cw(" + zval + ")";
  }

  // a * b, with a's real part only when a is known to be real
  static std::string mul(bool a_real, const std::string& a, const std::string& b) {
    return a_real ? "rmul(" + a + ".re, " + b + ")" : a + " * " + b;
  }

  // FP64 x_j for a residual: a register, or (xsplit) x0_j read from LDS ahead
  // of its row plus dx_j
  std::string xref(int j) {
    if (!xsplit) return "x" + std::to_string(j);
    std::string nm = "xl" + std::to_string(j);
    if (have_x.insert(j).second) loads.push_back("    const W " + nm + " = xs[" + std::to_string(64 * j) + " + lane];\n");
    return "(cz(" + nm + ") + cz(dx" + std::to_string(j) + "))";
  }

  // One factorisation + forward + backward pass in type F ("Z" or "W"),
  // appended to `out`. mode 0: right-hand side b; mode 1: scaled FP64
  // residual b - A x. Row i's back-substitution result t<i> is handed to
  // on_result.
  bool pass(const std::string& F, int mode) {
    const int n = P.n;
    have_g.clear();
    have_p.clear();
    have_x.clear();
    std::vector<std::string> code(n);
    std::vector<std::vector<std::string>> rload(n);
    std::vector<char> yz(n, 1);  // y'_i structurally zero
    // slot value has no imaginary part: its stamp is real and no elimination
    // step has updated it yet (products with it take the real form)
    std::vector<char> real(P.nnz, 0);
    for (int i = 0; i < n; i++) {
      std::ostringstream o;
      loads.clear();
      // gather row i
      for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++) {
        std::string re, im;
        if (!slot(s, re, im)) return false;
        real[s] = im.empty();
        if (mode == 1) {
          if (has_slot(s)) o << "    const Z a" << s << " = " << zexpr(re, im) << ";\n";
          o << "    " << F << " v" << s << " = "
            << (has_slot(s) ? conv(F, "a" + std::to_string(s)) : F + "{0, 0}") << ";\n";
        } else {
          o << "    " << F << " v" << s << " = " << conv(F, zexpr(re, im)) << ";\n";
        }
      }
      const bool gg = growth && mode == 0;
      if (gg) {  // ||A||_inf: this row's sum of |a|
        o << "    an = fmax(an, 0.0";
        for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++)
          if (has_slot(s)) o << " + c1(v" << s << ")";
        o << ");\n    gsr = 0.0;\n";
      }
      // right-hand side of row i
      bool ynz = false;
      std::string b;
      if (!rhs(i, b)) return false;
      if (mode == 0) {
        if (!b.empty()) {
          o << "    " << F << " y" << i << " = " << conv(F, "Z{" + b + ", 0.0}") << ";\n";
          if (track_bm) o << "    bm2 = fmax(bm2, m2(Z{" << b << ", 0.0}));\n";
          ynz = true;
        }
      } else {
        o << "    Z r" << i << " = " << (b.empty() ? std::string("Z{0.0, 0.0}") : "Z{" + b + ", 0.0}") << ";\n";
        for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++)
          if (has_slot(s))
            o << "    r" << i << " = r" << i << " - " << mul(real[s], "a" + std::to_string(s), xref(P.col[s])) << ";\n";
        o << "    " << F << " y" << i << " = " << conv(F, "scal(inv_s, r" + std::to_string(i) + ")") << ";\n";
        ynz = true;
      }
      // factor row i: the plan's factor program for this row
      const int32_t* fp = P.fprog + P.fprog_ptr[i];
      int nk = *fp++;
      for (int kk = 0; kk < nk; kk++) {
        int sik = fp[0], k = P.col[sik], nu = fp[2];
        fp += 3;
        // |l_ik| (row sum of |U| row k) <= |v_ik| w_k, v_ik final here
        if (gg) o << "    gsr = gsr + c1(v" << sik << ") * gw" << k << ";\n";
        for (int u = 0; u < nu; u++, fp += 2) {
          int dst = fp[0], src = fp[1];
          if (!(src > P.diag[k] && src < P.row_ptr[k + 1])) { why = "factor program reads outside U"; return false; }
          o << "    v" << dst << " = v" << dst << " - " << mul(real[sik], "v" + std::to_string(sik), "u" + std::to_string(src))
            << ";\n";
          real[dst] = 0;
        }
        if (!yz[k]) {
          const std::string pr = mul(real[sik], "v" + std::to_string(sik), "y" + std::to_string(k) + "s");
          if (!ynz) o << "    " << F << " y" << i << " = " << F << "{0, 0} - " << pr << ";\n";
          else o << "    y" << i << " = y" << i << " - " << pr << ";\n";
          ynz = true;
        }
      }
      // pivot: decomposition check as the generic kernels (|piv| > 0, finite);
      // refinement passes repeat the first pass' factorisation exactly, so
      // only the first pass checks every pivot; they repeat the check of
      // their last pivot, whose only effect is on register allocation (with
      // no check at all in the refinement passes, the mixed kernel of the
      // 50-section RC ladder takes 512 registers and 596 B of scratch per
      // lane instead of 360 registers)
      int d = P.diag[i];
      if (mode == 0 || i == n - 1) o <<"    { const Z pz = cz(v" << d << "); bad |= !(m2(pz) > 0.0) || !fin(pz); }\n";
      o << "    const " << F << " rd" << i << " = rcpF(v" << d << ");\n";
      if (gg) {  // row i of U: its sum of |u|, the diagonal term of |L||U|, and w_i = |1/u_ii| sum
        o << "    const double gr" << i << " = c1(v" << d << ")";
        for (int s = d + 1; s < P.row_ptr[i + 1]; s++) o << " + c1(v" << s << ")";
        o << ";\n    gm = fmax(gm, gsr + gr" << i << ");\n    const double gw" << i << " = c1(rd" << i << ") * gr" << i
          << ";\n";
      }
      for (int s = d + 1; s < P.row_ptr[i + 1]; s++)
        o << "    const " << F << " u" << s << " = " << mul(real[s], "v" + std::to_string(s), "rd" + std::to_string(i))
          << ";\n";
      if (ynz) {
        o << "    const " << F << " y" << i << "s = y" << i << " * rd" << i << ";\n";
        yz[i] = 0;
      }
      code[i] = o.str();
      rload[i] = loads;
    }
    // rows with their loads issued `lookahead` rows early
    std::ostringstream& O = out;
    auto issue = [&](int r) {
      if (r >= n || rload[r].empty()) return;
      for (auto& l : rload[r]) O << l;
      O << "    asm volatile(\"\" ::: \"memory\");\n";
    };
    for (int r = 0; r < lookahead; r++) issue(r);
    for (int i = 0; i < n; i++) {
      issue(i + lookahead);
      O << code[i];
      if (KG_LANE_FWD_BARRIER > 0 && (i + 1) % KG_LANE_FWD_BARRIER == 0) O << "    __builtin_amdgcn_sched_barrier(0);\n";
    }
    // back substitution
    for (int i = n - 1; i >= 0; i--) {
      std::string acc = yz[i] ? F + "{0, 0}" : "y" + std::to_string(i) + "s";
      O << "    " << F << " t" << i << " = " << acc << ";\n";
      for (int s = P.diag[i] + 1; s < P.row_ptr[i + 1]; s++)
        O << "    t" << i << " = t" << i << " - u" << s << " * t" << P.col[s] << ";\n";
      on_result(O, i);
      if (KG_LANE_BWD_BARRIER > 0 && i % KG_LANE_BWD_BARRIER == 0) O << "    __builtin_amdgcn_sched_barrier(0);\n";
    }
    return true;
  }
};

}  // namespace

const char* jit_ac_name(bool mixed, bool xsplit) {
  return !mixed ? "kg_ac_lane_c128" : xsplit ? "kg_ac_lane_c64_split" : "kg_ac_lane_c64";
}

bool jit_ac_source(const PatLayout& L, const int32_t* buf, bool mixed, bool xsplit, std::string& src,
                   std::string& why) {
  View P = view(L, buf);
  const int n = P.n;
  if (L.status != KG_OK) { why = "structurally singular pattern"; return false; }
  if (L.has_diode) { why = "diode small-signal stamps are not generated"; return false; }
  int nu = 0;
  for (int i = 0; i < n; i++) nu += P.row_ptr[i + 1] - P.diag[i] - 1;
  // registers per lane held across a pass: U' and y' (and the FP64 solution
  // for mixed precision); the register file holds 512 per lane at one wave
  // per SIMD, and the pass needs a few dozen more for its temporaries
  // (xsplit: dx in registers, x0 in 512 * n bytes of LDS per 64-lane block)
  int regs = !mixed ? 4 * (nu + n) : 2 * (nu + n) + (xsplit ? 2 * n : 4 * n);
  if (mixed && xsplit && n > 78) { why = "x0 would not fit in LDS at four waves per CU"; return false; }
  if (regs > 440) { why = "per-lane state exceeds the register budget"; return false; }
  Gen G(P);
  G.xsplit = mixed && xsplit;
  G.emajor = mixed;
  std::ostringstream& o = G.out;
  o << kPrelude;
  o << "extern \"C\" __global__ void __launch_bounds__(64) " << jit_ac_name(mixed, xsplit)
    << "(const double* __restrict__ prm,\n"
       "    const double* __restrict__ ebase, const double* __restrict__ freqs, int nf, i64 nsys,\n"
       "    double* __restrict__ out, int* __restrict__ status, double* __restrict__ err, int refine,\n"
       "    double stop, unsigned long long* __restrict__ fixq, i64 fixcap) {\n"
       "  const i64 sys0 = xcd_block() * 64 + threadIdx.x;\n";
  if (!mixed) {  // consecutive frequencies of one instance per wave (element loads shared by the wave)
    // lanes past the last system repeat it (identical values to the same
    // addresses; they queue nothing), so that the whole wave stays active for
    // the error figure's wave reduction
    o << "  const bool live = sys0 < nsys;\n  const i64 sys = live ? sys0 : nsys - 1;\n"
         "  const int inst = (int)(sys / nf), fi = (int)(sys - (i64)inst * nf);\n";
  } else {
    o << "  if (sys0 >= nsys) return;\n  const i64 sys = sys0;\n";
    // 64 instances at one frequency per wave: refinement converges alike
    // across the wave, so the adaptive stop applies to whole waves
    // (measured: 0.49 ms instead of 0.78 ms on the 4096 x 200 RC-ladder
    // sweep; the FP64 kernel is faster with the instance-major order)
    o << "  const i64 ni = nsys / nf, ib = sys / (64 * (i64)nf), rem = sys - ib * 64 * (i64)nf;\n"
         "  const int nbi = (int)(ni - ib * 64 < 64 ? ni - ib * 64 : 64);\n"
         "  const int fi = (int)(rem / nbi), inst = (int)(ib * 64 + rem % nbi);\n";
  }
  if (mixed) {
    // ib is uniform across the wave (64 * nf consecutive systems per group of
    // 64 instances): the group's block of values as a buffer resource
    // (descriptor word 3 of gfx9 raw buffers; reads past the block return 0)
    o << "  const __amdgpu_buffer_rsrc_t EBR = __builtin_amdgcn_make_buffer_rsrc((void*)(ebase + "
         "(i64)__builtin_amdgcn_readfirstlane((int)ib) * "
      << 64 * P.ne << "), (short)0, " << 512 * P.ne << ", 0x00020000);\n"
      << "  const int jl8 = (inst - (int)ib * 64) * 8;\n";
  }
  else o << "  const double* EB = ebase + (i64)inst * " << P.ne << ";\n";
  o << "  const double* PR = prm + (i64)inst * " << (int64_t)P.ne * KG_NPARAM << ";\n"
    << "  const double w = (2.0 * 3.14159265358979323846) * freqs[fi];\n"
    << "  Z* o = reinterpret_cast<Z*>(out) + ((i64)inst * nf + fi) * " << n << ";\n"
    << "  bool bad = false;\n";
  // growth guard: computed during the first pass' factorisation
  G.growth = true;
  int mrow = 1;
  for (int i = 0; i < n; i++) mrow = std::max(mrow, (int)(P.row_ptr[i + 1] - P.row_ptr[i]));
  o << "  double an = 0.0, gm = 0.0, gsr = 0.0;\n";
  if (!mixed) {
    G.on_result = [&](std::ostringstream& O, int i) { O << "    o[" << P.col_of[i] << "] = t" << i << ";\n"; };
    o << "  {\n";
    if (!G.pass("Z", 0)) { why = G.why; return false; }
    // rho = gm / an above KG_GUARD_GROWTH (or a failed pivot): re-solved by
    // the fixup; else the instance's figure is the backward-error bound
    // c_m rho of the solve
    // the figure is reduced over the wave first: its lanes hold one or two
    // instances (consecutive frequencies), and 64 atomics on one address
    // serialise (a segmented maximum over lanes of the same instance, then
    // one atomic per instance)
    o << "  }\n"
      << "  const bool queue = bad || !(gm <= " << fmt_double(KG_GUARD_GROWTH) << " * an);\n"
      << "  if (queue && live) fix_push(fixq, fixcap, (i64)inst * nf + fi, status, inst);\n"
      << "  if (err != nullptr) {\n"
      << "    const int lane = threadIdx.x;\n"
      << "    double ev = queue || !(gm > 0.0) ? 0.0 : " << fmt_double(guard_bound_coeff(mrow)) << " * (gm / an);\n"
      << "    for (int off = 1; off < 64; off <<= 1) {\n"
      << "      const double ov = __shfl_down(ev, off, 64);\n"
      << "      const int oi = __shfl_down(inst, off, 64);\n"
      << "      if (lane + off < 64 && oi == inst) ev = fmax(ev, ov);\n"
      << "    }\n"
      << "    const int pi = __shfl_up(inst, 1, 64);\n"
      << "    if ((lane == 0 || pi != inst) && ev > 0.0)\n"
      << "      atomicMax(reinterpret_cast<unsigned long long*>(&err[inst]), (unsigned long long)__double_as_longlong(ev));\n"
      << "  }\n"
      << "}\n";
    src = o.str();
    return true;
  } else {
    if (G.xsplit) {
      o << "  extern __shared__ W xs[];\n  const int lane = threadIdx.x;\n";
      for (int i = 0; i < n; i++) o << "  W dx" << i << " = W{0.0f, 0.0f};\n";
    } else {
      for (int i = 0; i < n; i++) o << "  Z x" << i << ";\n";
    }
    o << "  double est = 0.0, d2prev = 0.0, d2last = 0.0, bm2 = 0.0;\n  float xmf = 0.0f;\n  int npass = 0;\n";
    // pass 0: complex64 solve of b; x = its FP64 copy (xsplit: x0 = it, dx = 0)
    // |x|_inf is taken from this first solution (the refinement changes it by
    // the FP32 solve error only); tracking it through the refinement passes
    // makes the compiler keep far more values live
    G.on_result = [&](std::ostringstream& O, int i) {
      if (G.xsplit) O << "    xs[" << 64 * i << " + lane] = t" << i << "; xmf = cmax(xmf, t" << i << ");\n";
      else O << "    x" << i << " = cz(t" << i << "); xmf = cmax(xmf, t" << i << ");\n";
    };
    G.track_bm = true;
    o << "  {\n";
    if (!G.pass("W", 0)) { why = G.why; return false; }
    // the growth verdict is taken here, so that only a lane mask stays live
    // through the refinement
    o << "  }\n  bad |= !(gm <= " << fmt_double(KG_GUARD_GROWTH) << " * an);\n";
    G.track_bm = false;
    // refinement: corrections scaled by |b|_inf stay in FP32 range
    o << "  const double s = sqrt(bm2), inv_s = s > 0.0 ? 1.0 / s : 0.0, xm2 = (double)xmf * (double)xmf;\n"
         "  const float sf = (float)s;\n";
    // refinement passes after the first run inside a run-time loop: the pass
    // body is compiled once, so its registers are allocated once (an unrolled
    // sequence of passes lets live ranges of neighbouring passes overlap)
    G.on_result = [&](std::ostringstream& O, int i) {
      // (xsplit: the correction is formed and accumulated in complex64, the
      // precision dx is kept in; FP64 here costs registers, not accuracy)
      if (G.xsplit)
        O << "    { const W d = W{sf * t" << i << ".re, sf * t" << i << ".im}; dx" << i << " = W{dx" << i << ".re + d.re, dx"
          << i << ".im + d.im}; dmf = cmax(dmf, t" << i << "); }\n";
      else
        O << "    x" << i << " = x" << i << " + scal(s, cz(t" << i << ")); dmf = cmax(dmf, t" << i << ");\n";
    };
    // the first pass is peeled off the loop: a wave-uniform branch around the
    // remaining passes (adaptive stop) leaves the loop itself unchanged,
    // while an exit inside it makes the compiler spill
    o << "  if (s > 0.0 && refine > 0) {\n  {\n    float dmf = 0.0f;\n";
    G.lookahead = LOOKAHEAD_REFINE;
    if (!G.pass("W", 1)) { why = G.why; return false; }
    // only the squared correction sizes of the last two passes are carried;
    // the estimate is formed after the loop (square roots and divisions in
    // the loop body make the compiler keep far more values live)
    o << "    d2prev = d2last;\n    d2last = s * (double)dmf * (s * (double)dmf);\n    npass = 1;\n  }\n";
    // adaptive stop (stop > 0): a wave leaves after the first pass once every
    // lane's predicted remaining error, the squared relative correction times
    // KG_REFINE_SAFETY (refinement contracts by about the relative size of
    // its correction), is at most `stop`
    o << "  if (refine > 1 && !(stop > 0.0 && __builtin_amdgcn_ballot_w64(" << KG_REFINE_SAFETY
      << " * d2last <= stop * xm2) == __builtin_amdgcn_read_exec()))\n"
         "#pragma clang loop unroll(disable)\n"
         "  for (int it = 1; it < refine; it++) {\n    float dmf = 0.0f;\n";
    if (!G.pass("W", 1)) { why = G.why; return false; }
    o << "    d2prev = d2last;\n    d2last = s * (double)dmf * (s * (double)dmf);\n    npass = it + 1;\n  }\n  }\n"
      << "  {\n    const double dm = sqrt(d2last), dp = sqrt(d2prev), xm = sqrt(xm2);\n"
      << "    const double rho = (npass > 1 && dp > 0.0) ? fmin(dm / dp, 1.0) : 1.0;\n"
      << "    est = xm > 0.0 ? (npass == 1 && stop > 0.0 ? " << KG_REFINE_SAFETY
      << " * d2last / xm2 : rho * dm / xm) : 0.0;\n  }\n";
    for (int i = 0; i < n; i++) {
      if (G.xsplit) o << "  o[" << P.col_of[i] << "] = cz(xs[" << 64 * i << " + lane]) + cz(dx" << i << ");\n";
      else o << "  o[" << P.col_of[i] << "] = x" << i << ";\n";
    }
    // a solve whose complex64 factorisation grew (rho above
    // KG_GUARD_GROWTH), whose estimate stays above KG_GUARD_MIXED or whose
    // complex64 pivots failed is solved again in FP64 with partial pivoting
    // by the fixup
    o << "  if (bad || !(est <= " << fmt_double(KG_GUARD_MIXED) << "))\n    fix_push(fixq, fixcap, (i64)inst * nf + fi, status, inst);\n"
         "  else if (err != nullptr && est > 0.0)\n"
         "    atomicMax(reinterpret_cast<unsigned long long*>(&err[inst]), (unsigned long long)__double_as_longlong(est));\n";
  }
  o << "}\n";
  src = o.str();
  return true;
}

namespace {
std::mutex g_mu;
std::map<std::pair<int, std::string>, std::pair<hipModule_t, JitFn>> g_cache;
}  // namespace

hipError_t jit_load(int device, const std::string& src, const char* name, JitFn& out, std::string& log) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(device, src + "\n//" + name);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    out = it->second.second;
    return hipSuccess;
  }
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return e;
  std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "kg_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return hipErrorInvalidValue;
  }
  // packed FP32 instructions stay disabled, as for the prebuilt kernels
  const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  hiprtcResult rc = hiprtcCompileProgram(prog, 7, opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  if (rc != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return hipErrorInvalidValue;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  hipModule_t mod;
  if ((e = hipModuleLoadData(&mod, code.data())) != hipSuccess) return e;
  JitFn f;
  if ((e = hipModuleGetFunction(&f.fn, mod, name)) != hipSuccess) return e;
  (void)hipFuncGetAttribute(&f.vgprs, HIP_FUNC_ATTRIBUTE_NUM_REGS, f.fn);
  (void)hipFuncGetAttribute(&f.scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f.fn);
  g_cache[key] = {mod, f};
  out = f;
  return hipSuccess;
}

}  // namespace kg
