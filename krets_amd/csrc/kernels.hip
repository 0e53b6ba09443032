// gfx950 kernels of the MNA solve backend.
//
// Every system (one circuit instance at one analysis point) lives in LDS for
// its whole life: LU values in the pattern fixed by the host symbolic pass,
// the solution, the right-hand side and a work vector. TPS threads cooperate
// on one system (TPS = 1: lane per system, TPS = 64: wave per system, larger:
// workgroup per system) and SPB = blockDim/TPS systems share a block; all
// systems of a block share one pattern, so every phase is block-uniform and
// phases are separated by barriers. Rows are processed level by level:
//
//   stamp   per LU slot, sum the element contributions that land there
//           (krets' triplets + sum_triplets, stampable.rs / solver.rs:59-70)
//   factor  up-looking row LU without pivoting (pivots fixed on the host);
//           the reciprocal pivot is stored in the diagonal slot
//   solve   forward and backward substitution by dependency level
//   refine  (mixed precision) FP64 residual from freshly gathered FP64
//           stamps, scaled FP32 correction solve, FP64 update
//
// Analysis drivers: Newton for OP / DC sweep (solver/op.rs, solver/dc.rs),
// AC sweep (solver/ac.rs), fixed-step transient as one persistent launch
// (solver/transient.rs plus trapezoidal), LTE-controlled backward Euler and
// the discrete adjoint.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "kg_launch.h"

namespace kg {

// ------------------------------------------------------------ scalar types
template <class R>
struct cplx {
  R re, im;
};
using c64 = cplx<float>;
using c128 = cplx<double>;

template <class R>
__device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
__device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R>
__device__ __forceinline__ cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
template <class R>
__device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__device__ __forceinline__ cplx<R> operator*(R s, cplx<R> a) { return {s * a.re, s * a.im}; }
template <class R>
__device__ __forceinline__ cplx<R>& operator+=(cplx<R>& a, cplx<R> b) { a = a + b; return a; }
template <class R>
__device__ __forceinline__ cplx<R>& operator-=(cplx<R>& a, cplx<R> b) { a = a - b; return a; }

template <class S> struct Traits;
template <> struct Traits<double> { using R = double; static constexpr bool cx = false; };
template <> struct Traits<float> { using R = float; static constexpr bool cx = false; };
template <> struct Traits<c128> { using R = double; static constexpr bool cx = true; };
template <> struct Traits<c64> { using R = float; static constexpr bool cx = true; };

template <class S> __device__ __forceinline__ S mk(double re, double im);
template <> __device__ __forceinline__ double mk<double>(double re, double) { return re; }
template <> __device__ __forceinline__ float mk<float>(double re, double) { return (float)re; }
template <> __device__ __forceinline__ c128 mk<c128>(double re, double im) { return {re, im}; }
template <> __device__ __forceinline__ c64 mk<c64>(double re, double im) { return {(float)re, (float)im}; }

__device__ __forceinline__ double re_of(double a) { return a; }
__device__ __forceinline__ double re_of(float a) { return a; }
__device__ __forceinline__ double re_of(c128 a) { return a.re; }
__device__ __forceinline__ double re_of(c64 a) { return a.re; }
__device__ __forceinline__ double im_of(double) { return 0.0; }
__device__ __forceinline__ double im_of(float) { return 0.0; }
__device__ __forceinline__ double im_of(c128 a) { return a.im; }
__device__ __forceinline__ double im_of(c64 a) { return a.im; }

template <class D, class S>
__device__ __forceinline__ D cvt(S a) { return mk<D>(re_of(a), im_of(a)); }

// modulus used for pivot checks and norms
__device__ __forceinline__ double absv(double a) { return fabs(a); }
__device__ __forceinline__ double absv(float a) { return fabsf(a); }
__device__ __forceinline__ double absv(c128 a) { return sqrt(a.re * a.re + a.im * a.im); }
__device__ __forceinline__ double absv(c64 a) { return sqrt((double)a.re * a.re + (double)a.im * a.im); }

// reciprocal; complex via Smith's scaled division
__device__ __forceinline__ double rcp(double a) { return 1.0 / a; }
__device__ __forceinline__ float rcp(float a) { return 1.0f / a; }
template <class R>
__device__ __forceinline__ cplx<R> rcp(cplx<R> b) {
  if (fabs(b.re) >= fabs(b.im)) {
    R r = b.im / b.re, d = b.re + b.im * r;
    return {R(1) / d, -r / d};
  }
  R r = b.re / b.im, d = b.re * r + b.im;
  return {r / d, R(-1) / d};
}
__device__ __forceinline__ bool finite_v(double a) { return isfinite(a); }
__device__ __forceinline__ bool finite_v(float a) { return isfinite(a); }
template <class R>
__device__ __forceinline__ bool finite_v(cplx<R> a) { return isfinite(a.re) && isfinite(a.im); }
__device__ __forceinline__ double scal(double s, double a) { return s * a; }
__device__ __forceinline__ c128 scal(double s, c128 a) { return {s * a.re, s * a.im}; }

// ------------------------------------------------------- device maths
// constants.rs:1-4
__device__ constexpr double KG_VT = 1.380649e-23 * 300.0 / 1.602176634e-19;
__device__ constexpr double KG_PI = 3.14159265358979323846;

// Pulse::value_at / Sinusoidal::value_at (voltage_source.rs:50-93) and
// transient_value_at (243-252).
__device__ double wave_value(const double* p, double t) {
  int kind = (int)p[2];
  const double* w = p + 3;
  if (kind == KG_WAVE_PULSE) {
    double v1 = w[0], v2 = w[1], td = w[2], tr = w[3], tf = w[4], pw = w[5], per = w[6];
    if (t < td) return v1;
    double tt = fmod(t - td, per);
    if (tt < tr) return v1 + (v2 - v1) * (tt / tr);
    if (tt < tr + pw) return v2;
    if (tt < tr + pw + tf) return v2 + (v1 - v2) * ((tt - tr - pw) / tf);
    return v1;
  }
  if (kind == KG_WAVE_SIN) {
    double vo = w[0], va = w[1], f = w[2], td = w[3], th = w[4], ph = w[5];
    if (t < td) return vo;
    double om = 2.0 * KG_PI * f;
    double phr = ph * (KG_PI / 180.0);
    double tt = t - td;
    return vo + va * exp(-th * tt) * sin(om * tt + phr);
  }
  return p[0];
}

// Diode companion (elements/diode.rs:24-71). compat: krets' semantics, where
// the exponent limit is n*Vt*ln(DBL_MAX*n*Vt/Is) (infinite for any practical
// Is) and Ieq uses the unlimited voltage. clean: SPICE3 pnjlim on the
// junction voltage against the previous iterate, consistent linearisation.
struct DiodeLin {
  double g, ieq, vd;
  bool limited;
};
__device__ DiodeLin diode_eval(double vd, double is, double n, bool compat, double vold) {
  double nvt = n * KG_VT;
  DiodeLin r;
  r.limited = false;
  if (compat) {
    double vcrit = nvt * log(DBL_MAX * n * KG_VT / is);
    double vl = fmin(fmax(vd, -vcrit), vcrit);
    double ex = exp(vl / nvt);
    r.g = (is / nvt) * ex;
    double i = is * (ex - 1.0);
    r.ieq = i - r.g * vd;
    r.vd = vd;
    return r;
  }
  double vcrit = nvt * log(nvt / (1.4142135623730951 * is));
  double v = vd;
  if (v > vcrit && fabs(v - vold) > nvt + nvt) {
    if (vold > 0) {
      double arg = 1.0 + (v - vold) / nvt;
      v = arg > 0 ? vold + nvt * log(arg) : vcrit;
    } else {
      v = nvt * log(v / nvt);
    }
    r.limited = true;
  }
  double ex = exp(v / nvt);
  r.g = is / nvt * ex;
  r.ieq = is * (ex - 1.0) - r.g * v;
  r.vd = v;
  return r;
}

// Junction charge of a diode at junction voltage v (opt-in diode parasitics;
// SPICE2 model, Nagel 1975): depletion charge with parameters CJO, VJ, M, FC
// from the record pe ([2..5]),
//   q = CJO VJ / (1 - M) (1 - (1 - v/VJ)^(1-M)),  c = CJO (1 - v/VJ)^-M
// below FC VJ, continued linearly in c above it, plus the diffusion charge
// TT i (pe[6]) with i and its conductance g the diode's current at v.
struct JQ {
  double q, c;
};
__device__ __forceinline__ bool has_jq(const double* pe) { return pe[2] != 0.0 || pe[6] != 0.0; }
__device__ JQ junction_charge(double v, const double* pe, double i, double g) {
  const double cjo = pe[2], vj = pe[3], m = pe[4], fc = pe[5], tt = pe[6];
  JQ r{tt * i, tt * g};
  if (cjo != 0.0) {
    const double fcv = fc * vj;
    if (v < fcv) {
      const double a = 1.0 - v / vj, s = pow(a, -m);
      r.q += cjo * vj * (1.0 - a * s) / (1.0 - m);
      r.c += cjo * s;
    } else {
      const double f1 = vj * (1.0 - pow(1.0 - fc, 1.0 - m)) / (1.0 - m);
      const double f2 = pow(1.0 - fc, 1.0 + m), f3 = 1.0 - fc * (1.0 + m);
      r.q += cjo * (f1 + (f3 * (v - fcv) + m / (2.0 * vj) * (v * v - fcv * fcv)) / f2);
      r.c += cjo / f2 * (f3 + m * v / vj);
    }
  }
  return r;
}

// Diode current and conductance at v without limiting (junction charge of
// the previous step).
__device__ __forceinline__ void diode_iv(double v, double is, double n, double& i, double& g) {
  const double nvt = n * KG_VT, ex = exp(v / nvt);
  i = is * (ex - 1.0);
  g = is / nvt * ex;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2. Remapping gives every XCD a contiguous range of logical
// blocks, so systems that share an instance's element data (consecutive
// logical blocks) hit the same L2.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Largest block of a level-scheduled kernel with TPS threads per system:
// one wave (64 / TPS systems) below 64 threads per system, else one system.
// Declared as the kernels' launch bound, so that the register allocator
// budgets for it instead of for a 1024-thread block (128 VGPRs, which made
// the transient kernels spill about 300 bytes per lane).
#define KG_LB(TPS) __launch_bounds__((TPS) < 64 ? 64 : (TPS))

// Per-point stamp parameters.
struct StampArgs {
  int ak;            // AnalysisKind
  double freq;       // AC frequency (krets computes 2*PI*f*value left to right)
  double h;          // transient step
  double alpha;      // 1 = backward Euler, 2 = trapezoidal
  double tsrc;       // transient source time
  int sweep_e;       // DC-swept element or -1
  double sweep_val;  // its value at this point
};


template <class A>
__device__ __forceinline__ A elem_y(int e, const PatDev& P, const double* prm, const double* eb, const StampArgs& sa,
                                    const double* ed) {
  int ty = P.etype[e];
  if (ty == KG_R) return mk<A>(eb[e], 0.0);
  if (ty == KG_C) {
    double c = prm[e * KG_NPARAM];
    if (sa.ak == AK_AC) return mk<A>(0.0, 2.0 * KG_PI * sa.freq * c);
    return mk<A>(sa.alpha * (c / sa.h), 0.0);
  }
  if (ty == KG_D)  // AC: conductance and junction capacitance at the operating point
    return sa.ak == AK_AC ? mk<A>(ed[e * ESTRIDE], 2.0 * KG_PI * sa.freq * ed[e * ESTRIDE + 1])
                          : mk<A>(ed[e * ESTRIDE], 0.0);
  return mk<A>(0.0, 0.0);
}

template <class A>
__device__ __forceinline__ A elem_z(int e, const PatDev& P, const double* prm, const StampArgs& sa) {
  int ty = P.etype[e];
  double v = prm[e * KG_NPARAM];
  if (ty == KG_L) {
    if (sa.ak == AK_AC) return mk<A>(-0.0, -(2.0 * KG_PI * sa.freq * v));
    return mk<A>(-(sa.alpha * (v / sa.h)), 0.0);
  }
  if (ty == KG_R) return mk<A>(-v, 0.0);
  return mk<A>(1.0, 0.0);
}

template <class A>
__device__ __forceinline__ A mat_contrib(int32_t code, const PatDev& P, const double* prm, const double* eb,
                                         const StampArgs& sa, const double* ed) {
  int e = code >> 3, k = code & 7;
  switch (k) {
    case CK_YP: return elem_y<A>(e, P, prm, eb, sa, ed);
    case CK_YN: return -elem_y<A>(e, P, prm, eb, sa, ed);
    case CK_ONEP: return mk<A>(1.0, 0.0);
    case CK_ONEN: return mk<A>(-1.0, 0.0);
    case CK_KP: return mk<A>(eb[e], 0.0);
    case CK_KN: return mk<A>(-eb[e], 0.0);
    default: return elem_z<A>(e, P, prm, sa);
  }
}

// xp: previous-step solution (permuted numbering) for companion models.
template <class A>
__device__ __forceinline__ A rhs_contrib(int32_t code, const PatDev& P, const double* prm, const StampArgs& sa,
                                         const double* ed, const double* xp) {
  int e = code >> 3, k = code & 7;
  const double* pe = prm + e * KG_NPARAM;
  switch (k) {
    case RK_SRC: {
      if (e == sa.sweep_e) return mk<A>(sa.sweep_val, 0.0);
      if (P.etype[e] == KG_V) {
        if (sa.ak == AK_AC) return mk<A>(pe[1], 0.0);
        if (sa.ak == AK_TRAN) return mk<A>(wave_value(pe, sa.tsrc), 0.0);
      }
      return mk<A>(pe[0], 0.0);
    }
    case RK_DP: return mk<A>(-ed[e * ESTRIDE + 1], 0.0);
    case RK_DM: return mk<A>(ed[e * ESTRIDE + 1], 0.0);
    case RK_CP:
    case RK_CM: {
      int p = P.ep[e], m = P.em[e];
      double vprev = (p >= 0 ? xp[p] : 0.0) - (m >= 0 ? xp[m] : 0.0);
      // krets: i_eq = -(C/h) * v_prev; RHS[p] += -i_eq, RHS[m] += i_eq
      double ieq = -(sa.alpha * (pe[0] / sa.h)) * vprev;
      if (sa.alpha != 1.0) ieq -= ed[e * ESTRIDE];
      return mk<A>(k == RK_CP ? -ieq : ieq, 0.0);
    }
    default: {  // RK_IND: -(L/h) * i_prev (krets), minus v_prev for TRAP
      int c = P.ec[e];
      double r = -(sa.alpha * (pe[0] / sa.h)) * xp[c];
      if (sa.alpha != 1.0) r -= ed[e * ESTRIDE];
      return mk<A>(r, 0.0);
    }
  }
}

// ------------------------------------------------------------- phases
// All phase helpers are executed by every thread of the block; t is the
// thread's index within its system. Work is skipped when `act` is false but
// barriers are always reached.

template <class F, class A, int TPS>
__device__ void stamp_matrix(const PatDev& P, F* v, const double* prm, const double* eb, const StampArgs& sa,
                             const double* ed, int t, bool act) {
  if (act)
    for (int s = t; s < P.nnz; s += TPS) {
      A acc = mk<A>(0.0, 0.0);
      for (int q = P.scon_ptr[s]; q < P.scon_ptr[s + 1]; q++) acc += mat_contrib<A>(P.scon[q], P, prm, eb, sa, ed);
      v[s] = cvt<F>(acc);
    }
}

template <class A, int TPS>
__device__ void stamp_rhs(const PatDev& P, A* b, const double* prm, const StampArgs& sa, const double* ed,
                          const double* xp, int t, bool act) {
  if (act)
    for (int i = t; i < P.n; i += TPS) {
      A acc = mk<A>(0.0, 0.0);
      for (int q = P.rcon_ptr[i]; q < P.rcon_ptr[i + 1]; q++) acc += rhs_contrib<A>(P.rcon[q], P, prm, sa, ed, xp);
      b[i] = acc;
    }
}

// anorm (NORM, FP64): the largest row sum of |A| among the rows this thread
// factors, taken from each row just before its elimination (the static-pivot
// guard's ||A||_inf without re-gathering the stamps).
template <class F, int TPS, bool NORM = false>
__device__ void factor(const PatDev& P, F* v, int t, bool act, int* stat, double* anorm = nullptr) {
  auto row = [&](int i) {
    if constexpr (NORM) {
      double s = 0.0;
      for (int p = P.row_ptr[i]; p < P.row_ptr[i + 1]; p++) s += fabs(v[p]);
      *anorm = fmax(*anorm, s);
    }
    const int32_t* fp = P.fprog + P.fprog_ptr[i];
    int nk = *fp++;
    for (int kk = 0; kk < nk; kk++) {
      int sik = fp[0], sdk = fp[1], nu = fp[2];
      fp += 3;
      F lik = v[sik] * v[sdk];
      v[sik] = lik;
      for (int u = 0; u < nu; u++) {
        v[fp[0]] -= lik * v[fp[1]];
        fp += 2;
      }
    }
    F piv = v[P.diag[i]];
    double a = absv(piv);
    if (!(a > 0.0) || !finite_v(piv)) *stat = KG_DECOMPOSITION;
    v[P.diag[i]] = rcp(piv);
  };
  if constexpr (TPS == 1) {
    if (act)
      for (int i = 0; i < P.n; i++) row(i);
  } else {
    for (int l = 0; l < P.nlev; l++) {
      if (act)
        for (int q = P.lvl_ptr[l] + t; q < P.lvl_ptr[l + 1]; q += TPS) row(P.lvl_rows[q]);
      __syncthreads();
    }
  }
}

// Solves (LU) w = w in place.
template <class F, int TPS>
__device__ void lu_solve(const PatDev& P, const F* v, F* w, int t, bool act) {
  auto fwd = [&](int i) {
    F s = w[i];
    for (int p = P.row_ptr[i]; p < P.diag[i]; p++) s -= v[p] * w[P.col[p]];
    w[i] = s;
  };
  auto bwd = [&](int i) {
    F s = w[i];
    int d = P.diag[i];
    for (int p = d + 1; p < P.row_ptr[i + 1]; p++) s -= v[p] * w[P.col[p]];
    w[i] = s * v[d];
  };
  if constexpr (TPS == 1) {
    if (act) {
      for (int i = 0; i < P.n; i++) fwd(i);
      for (int i = P.n - 1; i >= 0; i--) bwd(i);
    }
  } else {
    for (int l = 0; l < P.nlev; l++) {
      if (act)
        for (int q = P.lvl_ptr[l] + t; q < P.lvl_ptr[l + 1]; q += TPS) fwd(P.lvl_rows[q]);
      __syncthreads();
    }
    for (int l = 0; l < P.nblev; l++) {
      if (act)
        for (int q = P.blvl_ptr[l] + t; q < P.blvl_ptr[l + 1]; q += TPS) bwd(P.blvl_rows[q]);
      __syncthreads();
    }
  }
}

// Solves (LU)^T w = w in place: U^T forward, then L^T backward.
template <class F, int TPS>
__device__ void lu_solve_t(const PatDev& P, const F* v, F* w, int t, bool act) {
  auto fwd = [&](int i) {
    F s = w[i];
    for (int q = P.ut_ptr[i]; q < P.ut_ptr[i + 1]; q += 2) s -= v[P.ut[q + 1]] * w[P.ut[q]];
    w[i] = s * v[P.diag[i]];
  };
  auto bwd = [&](int i) {
    F s = w[i];
    for (int q = P.lt_ptr[i]; q < P.lt_ptr[i + 1]; q += 2) s -= v[P.lt[q + 1]] * w[P.lt[q]];
    w[i] = s;
  };
  if constexpr (TPS == 1) {
    if (act) {
      for (int i = 0; i < P.n; i++) fwd(i);
      for (int i = P.n - 1; i >= 0; i--) bwd(i);
    }
  } else {
    for (int l = 0; l < P.ntlev; l++) {
      if (act)
        for (int q = P.tlvl_ptr[l] + t; q < P.tlvl_ptr[l + 1]; q += TPS) fwd(P.tlvl_rows[q]);
      __syncthreads();
    }
    for (int l = 0; l < P.ntblev; l++) {
      if (act)
        for (int q = P.tblvl_ptr[l] + t; q < P.tblvl_ptr[l + 1]; q += TPS) bwd(P.tblvl_rows[q]);
      __syncthreads();
    }
  }
}

// Max-reduction of a non-negative double over the TPS threads of a system.
template <int TPS>
__device__ double sys_max(double v, double* red, int ls, int t) {
  if constexpr (TPS == 1) {
    return v;
  } else {
    red[ls * TPS + t] = v;
    __syncthreads();
    for (int s = TPS / 2; s > 0; s >>= 1) {
      if (t < s) red[ls * TPS + t] = fmax(red[ls * TPS + t], red[ls * TPS + t + s]);
      __syncthreads();
    }
    double r = red[ls * TPS];
    __syncthreads();
    return r;
  }
}

// Full solve of A x = b given the factored v. FP64 (F == A): x := LU \ b.
// Mixed: x := LU \ b followed by `refine` steps of FP64 iterative refinement
// with A re-gathered in FP64. Returns the normwise size of the last
// correction relative to x (0 for FP64).
template <class F, class A, int TPS>
__device__ double solve_refine(const PatDev& P, const F* v, const A* b, A* x, F* w, const double* prm,
                               const double* eb, const StampArgs& sa, const double* ed, int refine, double* red,
                               int ls, int t, bool act, A* rbuf = nullptr) {
  if constexpr (std::is_same<F, A>::value) {
    if (act)
      for (int i = t; i < P.n; i += TPS) x[i] = b[i];
    __syncthreads();
    lu_solve<F, TPS>(P, v, x, t, act);
    return 0.0;
  } else {
    if (act)
      for (int i = t; i < P.n; i += TPS) w[i] = cvt<F>(b[i]);
    __syncthreads();
    lu_solve<F, TPS>(P, v, w, t, act);
    if (act)
      for (int i = t; i < P.n; i += TPS) x[i] = cvt<A>(w[i]);
    __syncthreads();
    double est = 0.0, dprev = 0.0;
    for (int it = 0; it < refine; it++) {
      // r = b - A x in FP64 (A gathered in FP64), kept in rbuf
      double rmax = 0.0;
      if (act)
        for (int i = t; i < P.n; i += TPS) {
          A r = b[i];
          for (int p = P.row_ptr[i]; p < P.row_ptr[i + 1]; p++) {
            int q0 = P.scon_ptr[p], q1 = P.scon_ptr[p + 1];
            if (q0 == q1) continue;
            A a = mk<A>(0.0, 0.0);
            for (int q = q0; q < q1; q++) a += mat_contrib<A>(P.scon[q], P, prm, eb, sa, ed);
            r -= a * x[P.col[p]];
          }
          rmax = fmax(rmax, absv(r));
          rbuf[i] = r;
        }
      double s = sys_max<TPS>(rmax, red, ls, t);
      // scaled FP32 correction right-hand side (keeps the residual in range)
      if (act && s > 0.0) {
        double inv = 1.0 / s;
        for (int i = t; i < P.n; i += TPS) w[i] = cvt<F>(scal(inv, rbuf[i]));
      }
      __syncthreads();
      lu_solve<F, TPS>(P, v, w, t, act && s > 0.0);  // barriers are reached by every system
      double dmax = 0.0, xmax = 0.0;
      if (act && s > 0.0)
        for (int i = t; i < P.n; i += TPS) {
          A d = scal(s, cvt<A>(w[i]));
          x[i] += d;
          dmax = fmax(dmax, absv(d));
          xmax = fmax(xmax, absv(x[i]));
        }
      __syncthreads();
      double dm = sys_max<TPS>(dmax, red, ls, t);
      double xm = sys_max<TPS>(xmax, red, ls, t);
      // remaining error ~ contraction ratio * last correction (geometric
      // convergence of refinement); the last correction itself bounds it
      double rho = (it > 0 && dprev > 0.0) ? fmin(dm / dprev, 1.0) : 1.0;
      est = xm > 0.0 ? rho * dm / xm : 0.0;
      dprev = dm;
    }
    return est;
  }
}

__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// Entry norm of the backward-error guard: |re| + |im|.
__device__ __forceinline__ double c1(double a) { return fabs(a); }
__device__ __forceinline__ double c1(c128 a) { return fabs(a.re) + fabs(a.im); }

// Componentwise backward error of an FP64 solve (Oettli-Prager), omega =
// max_i |b - A x|_i / ((|A| |x| + |b|)_i + u ||b||_inf) with |z| = |re| + |im|
// and A re-gathered in FP64 from the stamps (the system's v holds its
// factors); it bounds the normwise backward error. The floor u ||b||_inf
// (perturbations of b allowed down to FP64's resolution of its largest
// entry) keeps rows whose solution has decayed towards underflow measurable:
// a long RC ladder at high frequency attenuates its node voltages into the
// subnormal range, where the plain componentwise ratio is rounding noise of
// order 1 even for a partial-pivoting solve. Reached by every thread.
template <class A, int TPS>
__device__ double guard_omega(const PatDev& P, const A* b, const A* x, const double* prm, const double* eb,
                              const StampArgs& sa, const double* ed, double* red, int ls, int t, bool act) {
  double bm = 0.0;
  if (act)
    for (int i = t; i < P.n; i += TPS) bm = fmax(bm, c1(b[i]));
  const double floor_b = 0x1p-53 * sys_max<TPS>(bm, red, ls, t);
  double om = 0.0;
  if (act)
    for (int i = t; i < P.n; i += TPS) {
      A r = b[i];
      double dr = c1(b[i]);
      for (int p = P.row_ptr[i]; p < P.row_ptr[i + 1]; p++) {
        int q0 = P.scon_ptr[p], q1 = P.scon_ptr[p + 1];
        if (q0 == q1) continue;
        A a = mk<A>(0.0, 0.0);
        for (int q = q0; q < q1; q++) a += mat_contrib<A>(P.scon[q], P, prm, eb, sa, ed);
        const A xj = x[P.col[p]];
        r -= a * xj;
        dr += c1(a) * c1(xj);
      }
      const double ri = c1(r);
      const double den = dr + floor_b;
      if (ri > 0.0) om = fmax(om, den > 0.0 ? ri / den : INFINITY);
    }
  return sys_max<TPS>(om, red, ls, t);
}

// Normwise backward error |b - A x| / (|A| |x| + |b|) of a solve returned by
// mixed precision refinement (infinity norms over |re| + |im|, A re-gathered
// in FP64 from the stamps). Reached by every thread.
template <class A, int TPS>
__device__ double guard_eta(const PatDev& P, const A* b, const A* x, const double* prm, const double* eb,
                            const StampArgs& sa, const double* ed, double* red, int ls, int t, bool act) {
  double rm = 0.0, am = 0.0, bm = 0.0, xm = 0.0;
  if (act)
    for (int i = t; i < P.n; i += TPS) {
      A r = b[i];
      double ar = 0.0;
      for (int p = P.row_ptr[i]; p < P.row_ptr[i + 1]; p++) {
        int q0 = P.scon_ptr[p], q1 = P.scon_ptr[p + 1];
        if (q0 == q1) continue;
        A a = mk<A>(0.0, 0.0);
        for (int q = q0; q < q1; q++) a += mat_contrib<A>(P.scon[q], P, prm, eb, sa, ed);
        r -= a * x[P.col[p]];
        ar += c1(a);
      }
      rm = fmax(rm, c1(r));
      am = fmax(am, ar);
      bm = fmax(bm, c1(b[i]));
      xm = fmax(xm, c1(x[i]));
    }
  rm = sys_max<TPS>(rm, red, ls, t);
  am = sys_max<TPS>(am, red, ls, t);
  bm = sys_max<TPS>(bm, red, ls, t);
  xm = sys_max<TPS>(xm, red, ls, t);
  const double den = am * xm + bm;
  return den > 0.0 ? rm / den : (rm > 0.0 ? INFINITY : 0.0);
}

// Static-pivot guard of the FP64 OP, DC sweep and transient factorisations
// (KG_GUARD_STATIC): the lane AC kernels' a-priori bound guard_bound_coeff(m)
// * rho, rho = || |L||U| ||_inf / ||A||_inf, formed after the fact from the
// factors in v (multipliers below the diagonal, the pivot's reciprocal on it,
// the upper factor unscaled), with ||A||_inf taken by factor<..., NORM> as
// it eliminates. This thread's parts over the rows t, t + TPS, ...:
// pivot_guard_rows returns the longest row m and writes the row sums of |U|
// into us; after a barrier, pivot_guard_lu returns the rows of |L||U|.
template <int TPS>
__device__ double pivot_guard_rows(const PatDev& P, const double* v, double* us, int t, bool act) {
  double m = 0.0;
  if (act)
    for (int i = t; i < P.n; i += TPS) {
      const int d = P.diag[i];
      double u = 1.0 / fabs(v[d]);
      for (int p = d + 1; p < P.row_ptr[i + 1]; p++) u += fabs(v[p]);
      us[i] = u;
      m = fmax(m, (double)(P.row_ptr[i + 1] - P.row_ptr[i]));
    }
  return m;
}

template <int TPS>
__device__ double pivot_guard_lu(const PatDev& P, const double* v, const double* us, int t, bool act) {
  double gm = 0.0;
  if (act)
    for (int i = t; i < P.n; i += TPS) {
      double g = us[i];
      for (int p = P.row_ptr[i]; p < P.diag[i]; p++) g += fabs(v[p]) * us[P.col[p]];
      gm = fmax(gm, g);
    }
  return gm;
}

__device__ __forceinline__ double pivot_guard_bound(double gm, double an, double m) {
  return gm > 0.0 ? (an > 0.0 ? guard_bound_coeff((int)m) * (gm / an) : INFINITY) : 0.0;
}

// Maxima of three values over the TPS threads of a system: shuffles inside
// the wave (a system's threads are consecutive lanes), then, for systems of
// several waves, one exchange through red (TPS doubles of the system) between
// two barriers. Reached by every thread of the block.
template <int TPS>
__device__ void sys_max3(double& a, double& b, double& c, double* red, int ls, int t) {
  constexpr int W = TPS < 64 ? TPS : 64;
  for (int o = W / 2; o > 0; o >>= 1) {
    a = fmax(a, __shfl_xor(a, o, W));
    b = fmax(b, __shfl_xor(b, o, W));
    c = fmax(c, __shfl_xor(c, o, W));
  }
  if constexpr (TPS > 64) {
    double* r = red + ls * TPS;
    if ((t & 63) == 0) {
      r[3 * (t / 64)] = a;
      r[3 * (t / 64) + 1] = b;
      r[3 * (t / 64) + 2] = c;
    }
    __syncthreads();
    for (int k = 0; k < TPS / 64; k++) {
      a = fmax(a, r[3 * k]);
      b = fmax(b, r[3 * k + 1]);
      c = fmax(c, r[3 * k + 2]);
    }
    __syncthreads();
  }
}

// Level-scheduled form, reached by every thread of the system; an: this
// thread's part of ||A||_inf from factor<..., NORM>; us: n doubles of scratch.
template <int TPS>
__device__ double pivot_guard(const PatDev& P, const double* v, double an, double* us, double* red, int ls, int t,
                              bool act) {
  double m = pivot_guard_rows<TPS>(P, v, us, t, act);
  __syncthreads();
  double gm = pivot_guard_lu<TPS>(P, v, us, t, act);
  sys_max3<TPS>(an, m, gm, red, ls, t);
  return pivot_guard_bound(gm, an, m);
}

// Block-wide maximum of non-negative values into a 64-bit LDS word that was
// zeroed before a barrier: a wave maximum, then one atomic per wave.
__device__ __forceinline__ void block_max_to(double v, unsigned long long* w) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) atomicMax(w, (unsigned long long)__double_as_longlong(v));
}

// Queues system `sys` for k_ac_fix (queue words: [0] count, [1 + q] systems);
// a full queue marks the instance as failed instead.
__device__ __forceinline__ void fix_push(unsigned long long* fixq, int64_t cap, int64_t sys, int32_t* status,
                                         int inst) {
  const unsigned long long q = atomicAdd(fixq, 1ull);
  if (q < (unsigned long long)cap) fixq[1 + q] = (unsigned long long)sys;
  else atomicMax(&status[inst], KG_DECOMPOSITION);
}

// End of one AC solve (thread 0 of the system): a failed pivot, an FP64
// backward error above KG_GUARD_OMEGA or a mixed error estimate above
// KG_GUARD_MIXED queues the solve for k_ac_fix (when a queue is given);
// otherwise the status and the instance's error figure are published.
__device__ __forceinline__ void ac_finish(int32_t flag, double q, bool mixed, unsigned long long* fixq,
                                          int64_t fixcap, int64_t sys, int32_t* status, double* err, int inst) {
  const bool bad = flag != KG_OK || !(q <= (mixed ? KG_GUARD_MIXED : KG_GUARD_OMEGA));
  if (fixq && bad) {
    fix_push(fixq, fixcap, sys, status, inst);
    return;
  }
  if (flag != KG_OK) atomicMax(&status[inst], flag);
  if (err && q > 0.0) atomicMax(reinterpret_cast<unsigned long long*>(&err[inst]),
                                (unsigned long long)__double_as_longlong(q));
}

template <class F, class A>
struct SysMem {
  F* v;
  A* x;
  A* b;
  F* w;
  A* r;  // mixed precision: FP64 residual
  double* g;   // Newton guess (FP64)
  double* xp;  // previous solution(s) (transient); extra vectors follow at xp + k * xstride
  int xstride;
  double* ed;  // element state
  double* prm;  // staged element records: params [ne][KG_NPARAM], then the value table [ne]
  int* flags;  // [0] status [1] active [2] iterations [4] any diode limited
};

template <class F, class A>
__device__ SysMem<F, A> carve(char* base, const Lay& L) {
  SysMem<F, A> m;
  char* p = base;
  m.flags = (int*)p;
  p += 64;
  m.v = (F*)p;
  p += lay_al(sizeof(F) * (size_t)L.nnz_max);
  m.x = (A*)p;
  p += lay_al(sizeof(A) * (size_t)L.n_max);
  m.b = (A*)p;
  p += lay_al(sizeof(A) * (size_t)L.n_max);
  if (std::is_same<F, A>::value) {
    m.w = (F*)m.x;
    m.r = nullptr;
  } else {
    m.w = (F*)p;
    p += lay_al(sizeof(F) * (size_t)L.n_max);
    m.r = (A*)p;
    p += lay_al(sizeof(A) * (size_t)L.n_max);
  }
  m.g = L.need_g ? (double*)p : nullptr;
  if (L.need_g) p += lay_al(8 * (size_t)L.n_max);
  m.xp = L.need_xp ? (double*)p : nullptr;
  m.xstride = (int)(lay_al(8 * (size_t)L.n_max) / 8);
  p += (size_t)L.need_xp * lay_al(8 * (size_t)L.n_max);
  m.ed = L.need_ed ? (double*)p : nullptr;
  if (L.need_ed) p += lay_al(8 * (size_t)ESTRIDE * L.ne_max);
  m.prm = L.stage_prm ? (double*)p : nullptr;
  return m;
}

// Evaluates all diodes of a system at guess g (permuted numbering).
// compat_first: krets' empty solution map on the first OP iteration.
template <int TPS>
// jq (clean transients only; null elsewhere): the step (h, alpha) whose
// junction-charge companion is added to every diode that has one, with xp
// the previous step's solution: i_q = alpha (q(v) - q(v_prev)) / h -
// (alpha - 1) i_q,prev, linearised at the limited junction voltage.
__device__ void eval_diodes(const PatDev& P, const double* prm, const double* g, double* ed, bool compat,
                            bool compat_first, int* any_limited, int t, bool act, const StampArgs* jq = nullptr,
                            const double* xp = nullptr) {
  if (!act || !P.has_diode) return;
  for (int e = t; e < P.ne; e += TPS) {
    if (P.etype[e] != KG_D) continue;
    int p = P.ep[e], m = P.em[e];
    double vp, vm;
    if (compat) {  // diode.rs:26-36: absent key -> 0.5 / 0.0
      vp = (compat_first || p < 0) ? 0.5 : g[p];
      vm = (compat_first || m < 0) ? 0.0 : g[m];
    } else {
      vp = p >= 0 ? g[p] : 0.0;
      vm = m >= 0 ? g[m] : 0.0;
    }
    const double* pe = prm + e * KG_NPARAM;
    DiodeLin d = diode_eval(vp - vm, pe[0], pe[1], compat, ed[e * ESTRIDE + 2]);
    if (jq && has_jq(pe)) {
      const JQ q = junction_charge(d.vd, pe, d.ieq + d.g * d.vd, d.g);
      const double vo = (p >= 0 ? xp[p] : 0.0) - (m >= 0 ? xp[m] : 0.0);
      double io, go;
      diode_iv(vo, pe[0], pe[1], io, go);
      const JQ qo = junction_charge(vo, pe, io, go);
      const double gq = jq->alpha * q.c / jq->h;
      const double iq = jq->alpha * (q.q - qo.q) / jq->h - (jq->alpha - 1.0) * ed[e * ESTRIDE + 3];
      d.g += gq;
      d.ieq += iq - gq * d.vd;
    }
    ed[e * ESTRIDE] = d.g;
    ed[e * ESTRIDE + 1] = d.ieq;
    ed[e * ESTRIDE + 2] = d.vd;
    if (d.limited) *any_limited = 1;
  }
}

// Trapezoidal junction-charge currents after an accepted step from xo to xn
// (clean transients): i_q = 2 (q(v_n) - q(v_o)) / h - i_q,prev.
template <int TPS>
__device__ void update_junction_charge(const PatDev& P, const double* prm, const double* xo, const double* xn,
                                       double* ed, double h, int t, bool act) {
  if (!act || !P.has_diode) return;
  for (int e = t; e < P.ne; e += TPS) {
    if (P.etype[e] != KG_D) continue;
    const double* pe = prm + e * KG_NPARAM;
    if (!has_jq(pe)) continue;
    const int p = P.ep[e], m = P.em[e];
    const double vn = (p >= 0 ? xn[p] : 0.0) - (m >= 0 ? xn[m] : 0.0);
    const double vo = (p >= 0 ? xo[p] : 0.0) - (m >= 0 ? xo[m] : 0.0);
    double in, gn, io, go;
    diode_iv(vn, pe[0], pe[1], in, gn);
    diode_iv(vo, pe[0], pe[1], io, go);
    ed[e * ESTRIDE + 3] =
        2.0 * (junction_charge(vn, pe, in, gn).q - junction_charge(vo, pe, io, go).q) / h - ed[e * ESTRIDE + 3];
  }
}

// Newton convergence test. compat: solver.rs:76-104 (an empty previous map
// never converges; the bookkeeping key takes part with the voltage
// tolerance). xnew/gold are permuted-numbering solution and guess.
template <int TPS>
__device__ bool newton_converged(const PatDev& P, const double* xnew, const double* gold, double reltol,
                                 double iabstol, double vabstol, double* red, int ls, int t, bool act) {
  double bad = 0.0;
  if (act)
    for (int i = t; i < P.n; i += TPS) {
      double a = xnew[i], b = gold[i];
      double diff = fabs(a - b), scale = fmax(fabs(a), fabs(b));
      double atol = P.isbr[i] ? iabstol : vabstol;
      if (!(diff <= reltol * scale + atol)) bad = 1.0;
    }
  return sys_max<TPS>(bad, red, ls, t) == 0.0;
}

// ----------------------------------------------------------- kernels
// OP and DC sweep (solver/op.rs:15-86, solver/dc.rs:14-124); one system per
// instance, sweep points and Newton iterations inside the kernel.
// out: [instance][npts][n] (krets numbering). FP64.
template <int TPS>
__global__ void KG_LB(TPS) k_newton_dc(BatchDev B, int kind_slot, NewtonArgs na, int sweep_local_e, double start, double step,
                            int npts, double* out, const int64_t* out_off, int32_t* status, int32_t* iters,
                            Lay lay) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const int inst = (int)(lay.sys_off + xcd_block() * spb + ls);
  const bool valid = inst < B.n_instances;
  const PatDev& Pg = B.pats[B.inst_pat[kind_slot * B.n_instances + (valid ? inst : 0)]];
  size_t sb = sys_bytes(lay);
  char* base = sys_region(smem, lay);
  double* red = (double*)(base + spb * sb);
  SysMem<double, double> M = carve<double, double>(base + ls * sb, lay);
  // One large system per block (shared topology): the factor and solve
  // structure is copied into LDS behind the reduction buffer, so that the
  // dependent index loads of every level phase are LDS round trips rather
  // than L2 / MALL ones (the factorisation of a 1002-unknown ladder is a
  // chain of about 630 dependent updates).
  PatDev Ps = Pg;
  if (lay.stage_plan) {
    int4* dst = (int4*)(smem + spb * sb + 8 * (size_t)spb * TPS);
    const int4* src = (const int4*)Pg.base;
    for (int i = threadIdx.x; i < lay.stage_plan / 4; i += blockDim.x) dst[i] = src[i];
    const int32_t* pl = (const int32_t*)dst;
    auto rb = [&](const int32_t* p) { return p - Pg.base < lay.stage_plan ? pl + (p - Pg.base) : p; };
    Ps.row_ptr = rb(Pg.row_ptr); Ps.diag = rb(Pg.diag); Ps.fprog_ptr = rb(Pg.fprog_ptr); Ps.fprog = rb(Pg.fprog);
    Ps.lvl_ptr = rb(Pg.lvl_ptr); Ps.lvl_rows = rb(Pg.lvl_rows); Ps.blvl_ptr = rb(Pg.blvl_ptr);
    Ps.blvl_rows = rb(Pg.blvl_rows);
  }
  const PatDev& P = Ps;
  const double* prm = B.params + (valid ? B.eoff[inst] : 0) * KG_NPARAM;
  const double* eb = B.ebase + (valid ? B.eoff[inst] : 0);
  if (t == 0) {
    M.flags[0] = KG_OK;
    M.flags[1] = valid ? 1 : 0;
    M.flags[2] = 0;
  }
  __syncthreads();
    StampArgs sa{AK_DC, 0.0, 1.0, 1.0, 0.0, sweep_local_e, 0.0};
  bool have_last = false;  // dc.rs: warm start from the previous point
  for (int pt = 0; pt < npts; pt++) {
    sa.sweep_val = start + (double)pt * step;
    bool act = M.flags[1] != 0;
    bool prev_empty = na.compat ? !have_last : false;
    if (act && !na.compat && !have_last)
      for (int i = t; i < P.n; i += TPS) M.g[i] = 0.0;
    if (act && M.ed)  // element state exists when the batch has diodes
      for (int e = t; e < P.ne; e += TPS) M.ed[e * ESTRIDE + 2] = 0.0;
    __syncthreads();
    if (act && !na.compat && M.ed)
      for (int e = t; e < P.ne; e += TPS)
        if (P.etype[e] == KG_D) {
          int p = P.ep[e], m = P.em[e];
          M.ed[e * ESTRIDE + 2] = (p >= 0 ? M.g[p] : 0.0) - (m >= 0 ? M.g[m] : 0.0);
        }
    __syncthreads();
    int it = 0;
    bool done = !act;
    double anp = 0.0;  // this thread's part of ||A||_inf of the last factorisation
    for (it = 0; it < na.maxiter; it++) {
      bool a2 = !done;
      if (t == 0) M.flags[4] = 0;
      __syncthreads();
      eval_diodes<TPS>(P, prm, M.g, M.ed, na.compat != 0, prev_empty, &M.flags[4], t, a2);
      __syncthreads();
      stamp_matrix<double, double, TPS>(P, M.v, prm, eb, sa, M.ed, t, a2);
      stamp_rhs<double, TPS>(P, M.b, prm, sa, M.ed, nullptr, t, a2);
      __syncthreads();
      if (a2) anp = 0.0;  // a system that converged earlier keeps its last factorisation's
      factor<double, TPS, true>(P, M.v, t, a2, &M.flags[0], &anp);
      solve_refine<double, double, TPS>(P, M.v, M.b, M.x, M.w, prm, eb, sa, M.ed, 0, red, ls, t, a2);
      __syncthreads();
      bool stop;
      if (!P.nonlinear) {
        stop = true;
      } else if (na.compat && prev_empty) {
        stop = false;
        sys_max<TPS>(0.0, red, ls, t);
      } else {
        bool cv = newton_converged<TPS>(P, M.x, M.g, na.reltol, na.iabstol, na.vabstol, red, ls, t, a2);
        if (na.compat && it == 0 && pt > 0 && sweep_local_e >= 0) {
          // the "step" key: i vs i-1 on the first iteration of a point
          double v1 = (double)pt, v0 = (double)(pt - 1);
          cv = cv && fabs(v1 - v0) <= na.reltol * fmax(fabs(v1), fabs(v0)) + na.vabstol;
        }
        if (!na.compat && M.flags[4]) cv = false;
        stop = cv;
      }
      if (a2 && M.flags[0] != KG_OK) stop = true;
      // guess <- x for the next iteration
      if (a2 && !stop)
        for (int i = t; i < P.n; i += TPS) M.g[i] = M.x[i];
      prev_empty = false;
      __syncthreads();
      if (a2 && stop) {
        done = true;
        if (t == 0) M.flags[2] = it + 1;
      }
      int alldone = __syncthreads_and(done ? 1 : 0);
      if (alldone) break;
    }
    if (act && !done && t == 0) {
      M.flags[0] = KG_MAX_ITERATIONS;
      M.flags[2] = na.maxiter;
    }
    __syncthreads();
    {  // static-pivot guard on the factors of the point's solution (b is free: scratch)
      const bool g2 = act && M.flags[0] == KG_OK;
      const double bnd = pivot_guard<TPS>(P, M.v, anp, M.b, red, ls, t, g2);
      if (g2 && t == 0 && !(bnd <= KG_GUARD_STATIC)) M.flags[0] = KG_DECOMPOSITION;
      __syncthreads();
    }
    if (act) {
      double* o = out + out_off[inst] + (int64_t)pt * P.n;
      for (int j = t; j < P.n; j += TPS) o[P.col_of[j]] = M.x[j];
      for (int j = t; j < P.n; j += TPS) M.g[j] = M.x[j];
      if (t == 0 && iters) iters[(int64_t)inst * npts + pt] = M.flags[2];
    }
    have_last = true;
    __syncthreads();
    if (t == 0 && M.flags[0] != KG_OK) M.flags[1] = 0;
    __syncthreads();
  }
  if (valid && t == 0) status[inst] = M.flags[0];
}

// AC sweep (solver/ac.rs:17-109). One system per (instance, frequency).
// xop: operating points [instance][n] (krets numbering) for diode small-signal
// conductances. out: complex128 [instance][nf][n].
template <class F, class A, int TPS>
__global__ void KG_LB(TPS) k_ac(BatchDev B, int kind_slot, const double* freqs, int nf, const double* xop,
                     const int64_t* xop_off, double* out, const int64_t* out_off, int32_t* status, double* err,
                     int refine, Lay lay, unsigned long long* fixq, int64_t fixcap) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const int64_t sys = lay.sys_off + xcd_block() * spb + ls;
  const int64_t nsys = (int64_t)B.n_instances * nf;
  const bool valid = sys < nsys;
  const int inst = valid ? (int)(sys / nf) : 0;
  const int fi = valid ? (int)(sys % nf) : 0;
  const PatDev& P = B.pats[B.inst_pat[kind_slot * B.n_instances + inst]];
  size_t sb = sys_bytes(lay);
  char* base = sys_region(smem, lay);
  double* red = (double*)(base + spb * sb);
  SysMem<F, A> M = carve<F, A>(base + ls * sb, lay);
  const double* prm = B.params + B.eoff[inst] * KG_NPARAM;
  const double* eb = B.ebase + B.eoff[inst];
  if (t == 0) M.flags[0] = KG_OK;
  // diode small-signal conductance at the operating point (stampable.rs:614-641)
  if (valid && P.has_diode)
    for (int e = t; e < P.ne; e += TPS)
      if (P.etype[e] == KG_D) {
        const double* x0 = xop + xop_off[inst];
        int p = P.ep[e], m = P.em[e];
        double vp = p >= 0 ? x0[P.col_of[p]] : 0.5, vm = m >= 0 ? x0[P.col_of[m]] : 0.0;
        const double* pe = prm + e * KG_NPARAM;
        DiodeLin d = diode_eval(vp - vm, pe[0], pe[1], true, 0.0);
        M.ed[e * ESTRIDE] = d.g;
        M.ed[e * ESTRIDE + 1] = has_jq(pe) ? junction_charge(vp - vm, pe, d.ieq + d.g * (vp - vm), d.g).c : 0.0;
      }
  __syncthreads();
  StampArgs sa{AK_AC, valid ? freqs[fi] : 1.0, 1.0, 1.0, 0.0, -1, 0.0};
  stamp_matrix<F, A, TPS>(P, M.v, prm, eb, sa, M.ed, t, valid);
  stamp_rhs<A, TPS>(P, M.b, prm, sa, M.ed, nullptr, t, valid);
  __syncthreads();
  factor<F, TPS>(P, M.v, t, valid, &M.flags[0]);
  double est = solve_refine<F, A, TPS>(P, M.v, M.b, M.x, M.w, prm, eb, sa, M.ed, refine, red, ls, t, valid, M.r);
  __syncthreads();
  constexpr bool mixed = !std::is_same<F, A>::value;
  if constexpr (!mixed) {
    est = guard_omega<A, TPS>(P, M.b, M.x, prm, eb, sa, M.ed, red, ls, t, valid);
  } else if (!(guard_eta<A, TPS>(P, M.b, M.x, prm, eb, sa, M.ed, red, ls, t, valid) <= KG_GUARD_MIXED_ETA)) {
    est = INFINITY;  // the returned solution misses the normwise guard: to the fixup
  }
  if (valid) {
    double* o = out + out_off[inst] + (int64_t)fi * P.n * 2;
    for (int j = t; j < P.n; j += TPS) {
      int c = P.col_of[j];
      o[2 * c] = re_of(M.x[j]);
      o[2 * c + 1] = im_of(M.x[j]);
    }
    if (t == 0) ac_finish(M.flags[0], est, mixed, fixq, fixcap, sys, status, err, inst);
  }
}

// The plan with every pointer moved into a copy of its buffer at `lds`.
__device__ PatDev rebase(const PatDev& G, const int32_t* lds) {
  PatDev P = G;
  auto rb = [&](const int32_t* p) { return p ? lds + (p - G.base) : nullptr; };
  P.row_ptr = rb(G.row_ptr); P.col = rb(G.col); P.diag = rb(G.diag);
  P.fprog_ptr = rb(G.fprog_ptr); P.fprog = rb(G.fprog);
  P.lvl_ptr = rb(G.lvl_ptr); P.lvl_rows = rb(G.lvl_rows); P.blvl_ptr = rb(G.blvl_ptr); P.blvl_rows = rb(G.blvl_rows);
  P.scon_ptr = rb(G.scon_ptr); P.scon = rb(G.scon); P.rcon_ptr = rb(G.rcon_ptr); P.rcon = rb(G.rcon);
  P.col_of = rb(G.col_of); P.cpos = rb(G.cpos); P.isbr = rb(G.isbr);
  P.etype = rb(G.etype); P.ep = rb(G.ep); P.em = rb(G.em); P.ec = rb(G.ec);
  P.ut_ptr = rb(G.ut_ptr); P.ut = rb(G.ut); P.lt_ptr = rb(G.lt_ptr); P.lt = rb(G.lt);
  P.tlvl_ptr = rb(G.tlvl_ptr); P.tlvl_rows = rb(G.tlvl_rows); P.tblvl_ptr = rb(G.tblvl_ptr);
  P.tblvl_rows = rb(G.tblvl_rows);
  P.row_of = rb(G.row_of); P.rpos = rb(G.rpos); P.erp = rb(G.erp); P.erm = rb(G.erm); P.erc = rb(G.erc);
  P.base = lds;
  return P;
}

// Fixed-step transient (solver/transient.rs:14-114 and trapezoidal), one
// persistent launch: every instance runs all steps in-kernel. Linear
// circuits factor once. xop: OP per instance. out: [instance][nsteps+1][n].
// plan_ints > 0: the block's plan (all systems of a block share one) is
// copied into the first lay_al(4 * plan_ints) bytes of LDS once, so the
// thousands of steps read their row structure and level lists from LDS
// instead of issuing dependent L2 loads every step.
template <class F, class A, int TPS, bool STAGED>
__global__ void KG_LB(TPS) k_tran(BatchDev B, int kind_slot, NewtonArgs na, TranArgs ta, const double* xop,
                       const int64_t* xop_off, double* out, const int64_t* out_off, int32_t* status,
                       int32_t* iters, double* err, Lay lay, int plan_ints) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const int64_t slot0 = lay.sys_off + xcd_block() * spb;
  const int slot = (int)(slot0 + ls);
  const int nslot = B.inst_list ? B.n_list : B.n_instances;
  const bool valid = slot < nslot;
  const int inst = valid ? (B.inst_list ? B.inst_list[slot] : slot) : 0;
  const int ii = inst;
  char* sys_lds = sys_region(smem, lay);
  const PatDev& Pg = B.pats[B.inst_pat[kind_slot * B.n_instances + ii]];
  PatDev Ps;
  if constexpr (STAGED) {
    const int first = slot0 < nslot ? (B.inst_list ? B.inst_list[slot0] : (int)slot0) : 0;
    const PatDev& G = B.pats[B.inst_pat[kind_slot * B.n_instances + first]];
    int32_t* pl = (int32_t*)smem;
    for (int i = threadIdx.x; i < G.total; i += blockDim.x) pl[i] = G.base[i];
    __syncthreads();
    Ps = rebase(G, pl);
    if (!lay.gws) sys_lds = smem + lay_al(sizeof(int32_t) * (size_t)plan_ints);
  }
  const PatDev& P = [&]() -> const PatDev& {
    if constexpr (STAGED) return Ps;
    else return Pg;
  }();
  size_t sb = sys_bytes(lay);
  double* red = (double*)(sys_lds + spb * sb);
  SysMem<F, A> M = carve<F, A>(sys_lds + ls * sb, lay);
  const double* prm = B.params + B.eoff[ii] * KG_NPARAM;
  const double* eb = B.ebase + B.eoff[ii];
  const bool compat = ta.method == KG_BE_COMPAT;
  const bool trap = ta.method == KG_TRAP;
  const double* x0 = xop + xop_off[ii];
  if (t == 0) {
    M.flags[0] = valid ? status[inst] : KG_OK;
    M.flags[1] = (valid && M.flags[0] == KG_OK) ? 1 : 0;
  }
  __syncthreads();
  bool act = M.flags[1] != 0;
  // initial state: OP solution, capacitor currents 0, inductor voltages v(x0)
  if (act) {
    for (int j = t; j < P.n; j += TPS) M.xp[j] = x0[P.col_of[j]];
    double* o = out + out_off[ii];
    for (int j = t; j < P.n; j += TPS) o[j] = x0[j];
    if (t == 0 && iters) iters[(int64_t)ii * (ta.nsteps + 1)] = 0;
  }
  __syncthreads();
  if (act)
    for (int e = t; e < P.ne; e += TPS) {
      int p = P.ep[e], m = P.em[e];
      double vd = (p >= 0 ? M.xp[p] : 0.0) - (m >= 0 ? M.xp[m] : 0.0);
      M.ed[e * ESTRIDE] = P.etype[e] == KG_L ? vd : 0.0;
      M.ed[e * ESTRIDE + 2] = P.etype[e] == KG_D ? vd : 0.0;
      M.ed[e * ESTRIDE + 3] = 0.0;  // junction-charge current at the operating point
    }
  __syncthreads();
  StampArgs sa{AK_TRAN, 0.0, ta.h, ta.alpha, 0.0, -1, 0.0};
  const bool linear = !P.nonlinear;
  // FP64: the static-pivot guard's bound is the error figure (mixed: the refinement estimate)
  constexpr bool fp64 = std::is_same<F, double>::value;
  double est_max = 0.0;
  double anp = 0.0;  // FP64: this thread's part of ||A||_inf of the last factorisation
  if (linear) {
    stamp_matrix<F, A, TPS>(P, M.v, prm, eb, sa, M.ed, t, act);
    __syncthreads();
    factor<F, TPS, fp64>(P, M.v, t, act, &M.flags[0], &anp);
    __syncthreads();
    if constexpr (fp64) {  // static-pivot guard: one factorisation serves every step (b: scratch)
      const bool g2 = act && M.flags[0] == KG_OK;
      const double bnd = pivot_guard<TPS>(P, M.v, anp, M.b, red, ls, t, g2);
      if (g2) est_max = bnd;
      if (g2 && t == 0 && !(bnd <= KG_GUARD_STATIC)) M.flags[0] = KG_DECOMPOSITION;
      __syncthreads();
    }
  }
  for (int s = 1; s <= ta.nsteps; s++) {
    act = M.flags[1] != 0 && M.flags[0] == KG_OK;
    const double tn = (double)s * ta.h, tprev = (double)(s - 1) * ta.h;
    int used = 1;
    if (linear) {
      sa.tsrc = compat ? tprev : tn;  // krets evaluates sources at the guess's time
      stamp_rhs<A, TPS>(P, M.b, prm, sa, M.ed, M.xp, t, act);
      __syncthreads();
      double est =
          solve_refine<F, A, TPS>(P, M.v, M.b, M.x, M.w, prm, eb, sa, M.ed, ta.refine, red, ls, t, act, M.r);
      est_max = fmax(est_max, est);
    } else {
      if (act)
        for (int i = t; i < P.n; i += TPS) M.g[i] = M.xp[i];
      // clean Newton starts pnjlim from the junction voltages of the guess
      if (act && !compat)
        for (int e = t; e < P.ne; e += TPS)
          if (P.etype[e] == KG_D) {
            int p = P.ep[e], m = P.em[e];
            M.ed[e * ESTRIDE + 2] = (p >= 0 ? M.xp[p] : 0.0) - (m >= 0 ? M.xp[m] : 0.0);
          }
      __syncthreads();
      bool done = !act;
      int it;
      for (it = 0; it < na.maxiter; it++) {
        bool a2 = !done;
        sa.tsrc = (compat && it == 0) ? tprev : tn;
        if (t == 0) M.flags[4] = 0;
        __syncthreads();
        // junction charge in the clean integrators (krets' has none)
        eval_diodes<TPS>(P, prm, M.g, M.ed, compat, false, &M.flags[4], t, a2, compat ? nullptr : &sa, M.xp);
        __syncthreads();
        stamp_matrix<F, A, TPS>(P, M.v, prm, eb, sa, M.ed, t, a2);
        stamp_rhs<A, TPS>(P, M.b, prm, sa, M.ed, M.xp, t, a2);
        __syncthreads();
        if (a2) anp = 0.0;  // a system that converged earlier keeps its last factorisation's
        factor<F, TPS, fp64>(P, M.v, t, a2, &M.flags[0], &anp);
        double est =
            solve_refine<F, A, TPS>(P, M.v, M.b, M.x, M.w, prm, eb, sa, M.ed, ta.refine, red, ls, t, a2, M.r);
        est_max = fmax(est_max, est);
        __syncthreads();
        // convergence against the guess (FP64 copies)
        double bad = 0.0;
        if (a2)
          for (int i = t; i < P.n; i += TPS) {
            double a = re_of(M.x[i]), b = M.g[i];
            double diff = fabs(a - b), scale = fmax(fabs(a), fabs(b));
            double atol = P.isbr[i] ? na.iabstol : na.vabstol;
            if (!(diff <= na.reltol * scale + atol)) bad = 1.0;
          }
        bool cv = sys_max<TPS>(bad, red, ls, t) == 0.0;
        if (compat && it == 0)  // transient.rs:204, time key t_n vs t_{n-1}
          cv = cv && fabs(tn - tprev) <= na.reltol * fmax(fabs(tn), fabs(tprev)) + na.vabstol;
        if (!compat && M.flags[4]) cv = false;
        bool stop = cv || M.flags[0] != KG_OK;
        if (a2 && !stop)
          for (int i = t; i < P.n; i += TPS) M.g[i] = re_of(M.x[i]);
        __syncthreads();
        if (a2 && stop) {
          done = true;
          used = it + 1;
        }
        if (__syncthreads_and(done ? 1 : 0)) break;
      }
      if (act && !done) {
        if (t == 0) M.flags[0] = KG_MAX_ITERATIONS;
        used = na.maxiter;
      }
      __syncthreads();
      if constexpr (fp64) {  // static-pivot guard on the factors of the step's solution (b: scratch)
        const bool g2 = act && M.flags[0] == KG_OK;
        const double bnd = pivot_guard<TPS>(P, M.v, anp, M.b, red, ls, t, g2);
        if (g2) est_max = fmax(est_max, bnd);
        if (g2 && t == 0 && !(bnd <= KG_GUARD_STATIC)) M.flags[0] = KG_DECOMPOSITION;
        __syncthreads();
      }
      act = act && M.flags[0] == KG_OK;
    }
    // trapezoidal state update, store, advance
    if (act && trap)
      for (int e = t; e < P.ne; e += TPS) {
        int ty = P.etype[e];
        if (ty != KG_C && ty != KG_L) continue;
        int p = P.ep[e], m = P.em[e];
        double vn = (p >= 0 ? re_of(M.x[p]) : 0.0) - (m >= 0 ? re_of(M.x[m]) : 0.0);
        if (ty == KG_C) {
          double vo = (p >= 0 ? M.xp[p] : 0.0) - (m >= 0 ? M.xp[m] : 0.0);
          M.ed[e * ESTRIDE] = 2.0 * prm[e * KG_NPARAM] / ta.h * (vn - vo) - M.ed[e * ESTRIDE];
        } else {
          M.ed[e * ESTRIDE] = vn;
        }
      }
    if (trap) update_junction_charge<TPS>(P, prm, M.xp, M.x, M.ed, ta.h, t, act);  // x: FP64 in both precisions
    __syncthreads();
    if (act) {
      double* o = out + out_off[ii] + (int64_t)s * P.n;
      for (int j = t; j < P.n; j += TPS) {
        double v = re_of(M.x[j]);
        o[P.col_of[j]] = v;
        M.xp[j] = v;
      }
      if (t == 0 && iters) iters[(int64_t)ii * (ta.nsteps + 1) + s] = used;
    }
    __syncthreads();
  }
  if (valid && t == 0) {
    status[inst] = M.flags[0];
    if (err) err[inst] = est_max;
  }
}


// ------------------------------------------------------------ LDS-resident linear step loops
// One system per block of TPS threads, every thread owning at most R rows
// (n <= R * TPS). LDS holds the compact 16-bit plan (LinHdr), the LU factors,
// two solution buffers used alternately as x_n and x_{n-1}, and the
// coefficient and state of every capacitor and inductor companion; source
// parameters are copied in once. Each array ends in a zero slot (index n, nnz
// or ndyn) that padded work items point at, so that a thread's loads are
// unconditional. Source values of every step are tabulated up front in a
// global table (nsrc_max * (nsteps + 1) doubles per instance), so that no
// waveform code sits in a step loop; the thread that owns a source row loads
// each value one step ahead.
//
// The constant matrix is stamped and factored once from the full plan. Then
// every thread caches in registers what the plan says about its share of a
// step: its right-hand-side rows and their first LIN_KC contributions, its
// positions in the forward and backward level orders with their rows' first
// LIN_KE column indices, its companion elements and its columns in krets
// numbering. Within a phase a thread's LDS reads are independent: one round
// trip. Entries beyond the cached ones are read from the plan in LDS. The
// arithmetic is k_tran's on a linear circuit, operation for operation
// (solver/transient.rs:14-114 semantics, trapezoidal additions).
constexpr int LIN_KE = 4;  // cached strictly-triangular entries per row

// The entries of a long row beyond the cached ones, sum_p v[p] * x[col[p]]
// over [p, end): four independent partial sums, so the column and value
// loads of consecutive entries overlap instead of forming one dependent
// chain (the top separator rows of a nested-dissection order on a ladder
// hold up to ~2 log2 n entries, on the critical path of every solve).
__device__ __forceinline__ double row_tail(const double* v, const double* x, const uint16_t* col, int p, int end) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (; p + 4 <= end; p += 4) {
    const int c0 = col[p], c1 = col[p + 1], c2 = col[p + 2], c3 = col[p + 3];
    s0 += v[p] * x[c0];
    s1 += v[p + 1] * x[c1];
    s2 += v[p + 2] * x[c2];
    s3 += v[p + 3] * x[c3];
  }
  for (; p < end; p++) s0 += v[p] * x[col[p]];
  return (s0 + s1) + (s2 + s3);
}
constexpr int LIN_KC = 2;  // cached right-hand-side contributions per row

template <int TPS, int R>
struct LinSys {
  // plan views (LDS)
  const uint16_t *lw, *lvl_ptr, *blvl_ptr, *col, *cpos, *rcon_ptr, *rcon, *dyn, *src;
  const ushort4 *fdesc, *bdesc;
  int n, nnz, nlev, nblev, ndyn, nsrc, lev0, nsteps, nst1, t;
  // values (LDS), source table (global)
  double *v, *xc, *xq, *ed, *cf, *sp;
  double* tab;
  bool compat, trap, with_state;
  // register caches: right-hand side of row t + r*TPS (contribution = xq
  // slots a | b << 16, cf/ed slot d | flags << 16: bit 0 negate, bit 1
  // source, source index << 2); forward / backward positions t + r*TPS
  // (level, row | first << 16, count | diagonal << 16, column pairs);
  // companion element t + r*TPS (plus | minus << 16, kind); krets column
  // t + r*TPS: its permuted slot
  int rhs_n[R];
  uint32_t c_ab[R][LIN_KC], c_df[R][LIN_KC];
  int f_lev[R], b_lev[R];
  uint32_t f_ip[R], f_nd[R], f_col[R][LIN_KE / 2], b_ip[R], b_nd[R], b_col[R][LIN_KE / 2];
  uint32_t u_pm[R];
  int u_kind[R], o_slot[R];
  int pre_rc, pre_row;
  double pre;

  __device__ uint32_t node(uint16_t c) const { return c == LIN_NONE ? (uint32_t)n : (uint32_t)c; }

  // Plan copy (the first `words` words) and LDS carve-up; returns the first
  // free LDS byte. Every thread calls it; ends with a barrier.
  __device__ char* setup(const PatDev& P, int words, char* smem, const TranArgs& ta, double* srcv, int nsrc_max,
                         int inst) {
    t = threadIdx.x;
    {
      const uint4* s4 = reinterpret_cast<const uint4*>(P.lin);
      uint4* d4 = reinterpret_cast<uint4*>(smem);
      for (int i = t; i < words / 8; i += TPS) d4[i] = s4[i];
    }
    __syncthreads();
    const LinHdr* H = reinterpret_cast<const LinHdr*>(smem);
    lw = reinterpret_cast<const uint16_t*>(smem);
    n = H->n, nnz = H->nnz, nlev = H->nlev, nblev = H->nblev, ndyn = H->ndyn, nsrc = H->nsrc;
    lvl_ptr = lw + H->o_lvl_ptr;
    fdesc = reinterpret_cast<const ushort4*>(lw + H->o_fdesc);
    blvl_ptr = lw + H->o_blvl_ptr;
    bdesc = reinterpret_cast<const ushort4*>(lw + H->o_bdesc);
    col = lw + H->o_col;
    cpos = lw + H->o_cpos;
    rcon_ptr = lw + H->o_rcon_ptr;
    rcon = lw + H->o_rcon;
    dyn = lw + H->o_dyn;
    src = lw + H->o_src;
    lev0 = H->lev0_free ? 1 : 0;
    char* q = smem + lay_al(2 * (size_t)words);
    v = reinterpret_cast<double*>(q);
    q += 8 * ((size_t)nnz + 1);
    xc = reinterpret_cast<double*>(q);
    q += 8 * ((size_t)n + 1);
    xq = reinterpret_cast<double*>(q);
    q += 8 * ((size_t)n + 1);
    ed = reinterpret_cast<double*>(q);
    q += 8 * ((size_t)ndyn + 1);
    cf = reinterpret_cast<double*>(q);
    q += 8 * ((size_t)ndyn + 1);
    sp = reinterpret_cast<double*>(q);
    q += 8 * (size_t)KG_NPARAM * nsrc;
    nsteps = ta.nsteps;
    nst1 = ta.nsteps + 1;
    compat = ta.method == KG_BE_COMPAT;
    trap = ta.method == KG_TRAP;
    with_state = ta.alpha != 1.0;
    tab = srcv ? srcv + (int64_t)inst * nsrc_max * nst1 : nullptr;
    if (t == 0) {
      v[nnz] = 0.0;
      xc[n] = xq[n] = 0.0;
      ed[ndyn] = cf[ndyn] = 0.0;
    }
    return q;
  }

  // Stamps and factors the constant matrix (full plan P), fills the companion
  // coefficients, source records and the source table. flag: LDS status
  // word, followed by at least 28 free bytes. Returns the static-pivot
  // guard's bound (KG_GUARD_STATIC; a bound above it sets *flag).
  __device__ double factorise(const PatDev& P, const double* prm, const double* eb, const TranArgs& ta, int* flag) {
    StampArgs sa{AK_TRAN, 0.0, ta.h, ta.alpha, 0.0, -1, 0.0};
    unsigned long long* gw = reinterpret_cast<unsigned long long*>(flag + 2);  // ||A||, m, |||L||U|||
    if (t == 0) gw[0] = gw[1] = gw[2] = 0ull;
    stamp_matrix<double, double, TPS>(P, v, prm, eb, sa, ed, t, true);
    __syncthreads();
    double anp = 0.0;
    factor<double, TPS, true>(P, v, t, true, flag, &anp);
    for (int d = t; d < ndyn; d += TPS) {
      const double val = prm[(int)(dyn[d * LIN_DYN] & 0x3FFF) * KG_NPARAM];
      cf[d] = ta.alpha * (val / ta.h);
    }
    for (int i = t; i < nsrc * KG_NPARAM; i += TPS)
      sp[i] = prm[(int)src[(i / KG_NPARAM) * LIN_SRC + 1] * KG_NPARAM + i % KG_NPARAM];
    __syncthreads();
    // static-pivot guard (xc is free until the first step: scratch for the |U| row sums)
    block_max_to(anp, &gw[0]);
    block_max_to(pivot_guard_rows<TPS>(P, v, xc, t, true), &gw[1]);
    __syncthreads();
    block_max_to(pivot_guard_lu<TPS>(P, v, xc, t, true), &gw[2]);
    __syncthreads();
    const double bnd = pivot_guard_bound(__longlong_as_double((long long)gw[2]), __longlong_as_double((long long)gw[0]),
                                         __longlong_as_double((long long)gw[1]));
    if (t == 0 && *flag == KG_OK && !(bnd <= KG_GUARD_STATIC)) *flag = KG_DECOMPOSITION;
    for (int i = t; i < nsrc * nst1; i += TPS) {
      const int sidx = i / nst1, s = i - sidx * nst1;
      const double tn = (double)s * ta.h, tprev = (double)(s - 1) * ta.h;
      const double tsrc = compat ? tprev : tn;  // krets evaluates sources at the guess's time
      const double* pe = sp + sidx * KG_NPARAM;
      tab[i] = s == 0 ? 0.0 : (src[sidx * LIN_SRC] == KG_V ? wave_value(pe, tsrc) : pe[0]);
    }
    return bnd;
  }

  __device__ void contrib(int code, uint32_t& ab, uint32_t& df) const {
    const uint32_t NN = (uint32_t)n;
    const int idx = code >> 2, k = code & 3;
    if (k == LK_SRC) {
      ab = NN | (NN << 16);
      df = (uint32_t)ndyn | ((2u | ((uint32_t)idx << 2)) << 16);
    } else if (k == LK_IND) {
      ab = node(dyn[idx * LIN_DYN + 3]) | (NN << 16);
      df = (uint32_t)idx;
    } else {
      ab = node(dyn[idx * LIN_DYN + 1]) | (node(dyn[idx * LIN_DYN + 2]) << 16);
      df = (uint32_t)idx | ((k == LK_CP ? 1u : 0u) << 16);
    }
  }

  // Fills the register caches; the caller puts a barrier before the first
  // step (the source table must be complete).
  __device__ void cache() {
    const uint32_t NN = (uint32_t)n;
    auto level_of = [&](const uint16_t* ptr, int nl, int g) {
      int l = 0;
      while (l + 1 < nl && (int)ptr[l + 1] <= g) l++;
      return l;
    };
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int g = t + r * TPS;
      const bool ok = g < n;
      const int k0 = ok ? rcon_ptr[g] : 0, k1 = ok ? rcon_ptr[g + 1] : 0;
      rhs_n[r] = ok ? k1 - k0 : -1;
#pragma unroll
      for (int c = 0; c < LIN_KC; c++) {
        c_ab[r][c] = NN | (NN << 16);
        c_df[r][c] = (uint32_t)ndyn;
        if (k0 + c < k1) contrib(rcon[k0 + c], c_ab[r][c], c_df[r][c]);
      }
      f_lev[r] = b_lev[r] = -1;
      f_ip[r] = f_nd[r] = b_ip[r] = b_nd[r] = 0;
#pragma unroll
      for (int k = 0; k < LIN_KE / 2; k++) f_col[r][k] = b_col[r][k] = NN | (NN << 16);
      if (ok) {
        const ushort4 fd = fdesc[g], bd = bdesc[g];
        f_lev[r] = level_of(lvl_ptr, nlev, g);
        f_ip[r] = (uint32_t)fd.x | ((uint32_t)fd.y << 16);
        f_nd[r] = (uint32_t)(fd.z - fd.y) | ((uint32_t)fd.w << 16);
        b_lev[r] = level_of(blvl_ptr, nblev, g);
        b_ip[r] = (uint32_t)bd.x | ((uint32_t)bd.y << 16);
        b_nd[r] = (uint32_t)(bd.z - bd.y) | ((uint32_t)bd.w << 16);
#pragma unroll
        for (int k = 0; k < LIN_KE; k += 2) {
          const int fc = fd.z - fd.y, bc = bd.z - bd.y;
          const uint32_t c0 = k < fc ? col[fd.y + k] : NN, c1 = k + 1 < fc ? col[fd.y + k + 1] : NN;
          f_col[r][k / 2] = c0 | (c1 << 16);
          const uint32_t e0 = k < bc ? col[bd.y + k] : NN, e1 = k + 1 < bc ? col[bd.y + k + 1] : NN;
          b_col[r][k / 2] = e0 | (e1 << 16);
        }
      }
      u_kind[r] = g < ndyn ? (int)(dyn[g * LIN_DYN] >> 14) : 0;
      u_pm[r] = g < ndyn ? node(dyn[g * LIN_DYN + 1]) | (node(dyn[g * LIN_DYN + 2]) << 16) : (NN | (NN << 16));
      o_slot[r] = ok ? (int)cpos[g] : (int)NN;
    }
    pre_rc = -1;
    pre_row = 0;
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int c = 0; c < LIN_KC; c++)
        if (pre_rc < 0 && ((c_df[r][c] >> 16) & 2u)) {
          pre_rc = r * LIN_KC + c;
          pre_row = (int)(c_df[r][c] >> 18) * nst1;
        }
    pre = 0.0;
  }

  // first prefetch of the source value of step s (after the table barrier)
  __device__ void prefetch(int s) {
    if (pre_rc >= 0 && s <= nsteps) pre = tab[pre_row + s];
  }

  __device__ double rhs_one(int code, int s) const {
    uint32_t ab, df;
    contrib(code, ab, df);
    if ((df >> 16) & 2u) return tab[(int)(df >> 18) * nst1 + s];
    double ieq = -cf[df & 0xFFFF] * (xq[ab & 0xFFFF] - xq[ab >> 16]);
    if (with_state) ieq -= ed[df & 0xFFFF];
    return (df >> 16) & 1u ? -ieq : ieq;
  }

  // Right-hand side of step s into xc (permuted rows); prefetches step s+1.
  __device__ void rhs(int s) {
    const double pre_now = pre;
    if (pre_rc >= 0 && s < nsteps) pre = tab[pre_row + s + 1];
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (rhs_n[r] < 0) continue;
      const int i = t + r * TPS;
      double xa[LIN_KC], xb[LIN_KC], cv[LIN_KC], ev[LIN_KC];
#pragma unroll
      for (int c = 0; c < LIN_KC; c++) {
        xa[c] = xq[c_ab[r][c] & 0xFFFF];
        xb[c] = xq[c_ab[r][c] >> 16];
        cv[c] = cf[c_df[r][c] & 0xFFFF];
        ev[c] = ed[c_df[r][c] & 0xFFFF];
      }
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < LIN_KC; c++) {
        const uint32_t fl = c_df[r][c] >> 16;
        double val;
        if (fl & 2u) {
          val = r * LIN_KC + c == pre_rc ? pre_now : tab[(int)(fl >> 2) * nst1 + s];
        } else {
          double ieq = -cv[c] * (xa[c] - xb[c]);
          if (with_state) ieq -= ev[c];
          val = fl & 1u ? -ieq : ieq;
        }
        acc += val;  // padded contributions add -0.0: no change
      }
      if (rhs_n[r] > LIN_KC)
        for (int k = rcon_ptr[i] + LIN_KC; k < rcon_ptr[i + 1]; k++) acc += rhs_one(rcon[k], s);
      xc[i] = acc;
    }
  }

  // L y = b, U x = y in place in xc; ends with a barrier. Level 0 rows, when
  // free of lower entries, are the right-hand side itself.
  __device__ void solve() {
    for (int l = lev0; l < nlev; l++) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (f_lev[r] != l) continue;
        const int i = (int)(f_ip[r] & 0xFFFF), p0 = (int)(f_ip[r] >> 16), cnt = (int)(f_nd[r] & 0xFFFF);
        double sum = xc[i];
        double vv[LIN_KE], xx[LIN_KE];
#pragma unroll
        for (int k = 0; k < LIN_KE; k++) {
          vv[k] = v[k < cnt ? p0 + k : nnz];
          xx[k] = xc[k & 1 ? f_col[r][k / 2] >> 16 : f_col[r][k / 2] & 0xFFFF];
        }
#pragma unroll
        for (int k = 0; k < LIN_KE; k++) sum -= vv[k] * xx[k];  // padded entries subtract 0 * 0
        if (cnt > LIN_KE) sum -= row_tail(v, xc, col, p0 + LIN_KE, p0 + cnt);
        xc[i] = sum;
      }
      __syncthreads();
    }
    for (int l = 0; l < nblev; l++) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (b_lev[r] != l) continue;
        const int i = (int)(b_ip[r] & 0xFFFF), p0 = (int)(b_ip[r] >> 16), cnt = (int)(b_nd[r] & 0xFFFF);
        const int d = (int)(b_nd[r] >> 16);
        double sum = xc[i];
        const double vd = v[d];
        double vv[LIN_KE], xx[LIN_KE];
#pragma unroll
        for (int k = 0; k < LIN_KE; k++) {
          vv[k] = v[k < cnt ? p0 + k : nnz];
          xx[k] = xc[k & 1 ? b_col[r][k / 2] >> 16 : b_col[r][k / 2] & 0xFFFF];
        }
#pragma unroll
        for (int k = 0; k < LIN_KE; k++) sum -= vv[k] * xx[k];
        if (cnt > LIN_KE) sum -= row_tail(v, xc, col, p0 + LIN_KE, p0 + cnt);
        xc[i] = sum * vd;  // diagonal stored as its reciprocal
      }
      __syncthreads();
    }
  }

  // Trapezoidal companion states from x_n (xc) and x_{n-1} (xq).
  __device__ void update() {
    if (!trap) return;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int d = t + r * TPS;
      if (d >= ndyn) continue;
      const uint32_t a = u_pm[r] & 0xFFFF, b = u_pm[r] >> 16;
      const double vn = xc[a] - xc[b], vo = xq[a] - xq[b], c = cf[d], e = ed[d];
      ed[d] = u_kind[r] == 1 ? c * (vn - vo) - e : vn;  // 2C/h == alpha * C / h for alpha = 2
    }
    for (int d = t + R * TPS; d < ndyn; d += TPS) {  // companions beyond the cached ones
      const uint16_t* r = dyn + d * LIN_DYN;
      const double vn = xc[node(r[1])] - xc[node(r[2])];
      ed[d] = (r[0] >> 14) == 1 ? cf[d] * (vn - (xq[node(r[1])] - xq[node(r[2])])) - ed[d] : vn;
    }
  }

  __device__ void swap() {
    double* sw = xc;
    xc = xq;
    xq = sw;
  }

  // initial companion states from the state in xq: inductors at v(x0),
  // capacitors at 0 A
  __device__ void init_state() {
    for (int d = t; d < ndyn; d += TPS) {
      const uint16_t* r = dyn + d * LIN_DYN;
      ed[d] = (r[0] >> 14) == 2 ? xq[node(r[1])] - xq[node(r[2])] : 0.0;
    }
  }
};

// Fixed-step linear transient (k_tran semantics) from LDS; see LinSys.
template <int TPS, int R>
__global__ void __launch_bounds__(TPS) k_tran_lin(BatchDev B, int kind_slot, TranArgs ta, const double* xop,
                                                  const int64_t* xop_off, double* out, const int64_t* out_off,
                                                  int32_t* status, int32_t* iters, double* err, double* srcv,
                                                  int nsrc_max) {
  extern __shared__ __align__(16) char smem[];
  const int t = threadIdx.x;
  const int nslot = B.inst_list ? B.n_list : B.n_instances;
  const int64_t slot = xcd_block();
  if (slot >= nslot) return;
  const int inst = B.inst_list ? B.inst_list[slot] : (int)slot;
  const PatDev& P = B.pats[B.inst_pat[kind_slot * B.n_instances + inst]];
  if (status[inst] != KG_OK) {  // failed operating point or no plan: nothing to integrate
    if (t == 0 && err) err[inst] = 0.0;
    return;
  }
  if (P.lin == nullptr || P.n > R * TPS || (srcv == nullptr && nsrc_max > 0)) {  // launched only where it fits
    if (t == 0) status[inst] = KG_UNEXPECTED;
    return;
  }
  LinSys<TPS, R> S;
  const int words = reinterpret_cast<const LinHdr*>(P.lin)->words_fwd;
  int* flags = reinterpret_cast<int*>(S.setup(P, words, smem, ta, srcv, nsrc_max, inst));
  if (t == 0) flags[0] = KG_OK;
  __syncthreads();
  const double bnd = S.factorise(P, B.params + B.eoff[inst] * KG_NPARAM, B.ebase + B.eoff[inst], ta, &flags[0]);
  const double* x0 = xop + xop_off[inst];
  double* o = out + out_off[inst];
  for (int j = t; j < S.n; j += TPS) {
    const double xv = x0[j];
    S.xq[S.cpos[j]] = xv;
    o[j] = xv;
  }
  if (t == 0 && iters) iters[(int64_t)inst * (ta.nsteps + 1)] = 0;
  __syncthreads();
  if (flags[0] != KG_OK) {
    if (t == 0) {
      status[inst] = flags[0];
      if (err) err[inst] = bnd;
    }
    return;
  }
  S.init_state();
  S.cache();
  __syncthreads();  // also publishes the source table within the block
  S.prefetch(1);
  for (int s = 1; s <= ta.nsteps; s++) {
    S.rhs(s);
    __syncthreads();
    S.solve();
    S.update();
    double* os = o + (int64_t)s * S.n;
#pragma unroll
    for (int r = 0; r < R; r++)
      if (t + r * TPS < S.n) os[t + r * TPS] = S.xc[S.o_slot[r]];
    if (t == 0 && iters) iters[(int64_t)inst * (ta.nsteps + 1) + s] = 1;
    __syncthreads();
    S.swap();
  }
  if (t == 0) {
    status[inst] = flags[0];
    if (err) err[inst] = bnd;
  }
}

// Block sum of one double per thread (buf: TPS / 64 doubles of LDS).
template <int TPS>
__device__ double lin_block_sum(double v, double* buf) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (TPS == 64) return v;
  if ((threadIdx.x & 63) == 0) buf[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < TPS / 64; w++) s += buf[w];
  __syncthreads();
  return s;
}

// Forward pass of the discrete adjoint from LDS (linear circuits, clean BE
// or trapezoidal; k_adj_fwd semantics): the transient, the objective
// J = sum_s w_s c.x_s (trapezoid weights) and a checkpoint every K steps of
// the state: the solution (permuted order) then the companion states.
// ckpt: [instance][ncp][cpw], cpw >= n + ndyn; cvec in krets numbering.
template <int TPS, int R>
__global__ void __launch_bounds__(TPS) k_adj_lin_fwd(BatchDev B, int kind_slot, TranArgs ta, int K,
                                                     const double* cvec, const double* xop, const int64_t* xop_off,
                                                     double* ckpt, int64_t ck_stride, int cpw, double* Jout,
                                                     int32_t* status, double* srcv, int nsrc_max) {
  extern __shared__ __align__(16) char smem[];
  const int t = threadIdx.x;
  const int inst = (int)xcd_block();
  if (inst >= B.n_instances) return;
  const PatDev& P = B.pats[B.inst_pat[kind_slot * B.n_instances + inst]];
  if (status[inst] != KG_OK) return;
  if (P.lin == nullptr || P.n > R * TPS || (srcv == nullptr && nsrc_max > 0)) {
    if (t == 0) status[inst] = KG_UNEXPECTED;
    return;
  }
  LinSys<TPS, R> S;
  const LinHdr* Hg = reinterpret_cast<const LinHdr*>(P.lin);
  char* q = S.setup(P, Hg->words, smem, ta, srcv, nsrc_max, inst);
  q += 8 * (4 * ((size_t)S.n + 1) + (size_t)S.ndyn + 1 + (size_t)Hg->ne + 1);  // backward-only arrays
  double* red = reinterpret_cast<double*>(q);
  int* flags = reinterpret_cast<int*>(q + 8 * 32);
  if (t == 0) flags[0] = KG_OK;
  __syncthreads();
  S.factorise(P, B.params + B.eoff[inst] * KG_NPARAM, B.ebase + B.eoff[inst], ta, &flags[0]);
  const double* x0 = xop + xop_off[inst];
  for (int j = t; j < S.n; j += TPS) S.xq[S.cpos[j]] = x0[j];
  __syncthreads();
  if (flags[0] != KG_OK) {
    if (t == 0) status[inst] = flags[0];
    return;
  }
  S.init_state();
  S.cache();
  double cv[R];
#pragma unroll
  for (int r = 0; r < R; r++) cv[r] = t + r * TPS < S.n ? cvec[t + r * TPS] : 0.0;
  __syncthreads();
  S.prefetch(1);
  double* ck = ckpt + (int64_t)inst * ck_stride;
  double jpart = 0.0;
  for (int s = 0; s <= ta.nsteps; s++) {
    if (s > 0) {
      S.rhs(s);
      __syncthreads();
      S.solve();
      S.update();
      __syncthreads();
      S.swap();  // xq: x_s
    }
    const double w = (s == 0 || s == ta.nsteps) ? 0.5 * ta.h : ta.h;
#pragma unroll
    for (int r = 0; r < R; r++)
      if (t + r * TPS < S.n) jpart += w * cv[r] * S.xq[S.o_slot[r]];
    if (s % K == 0) {
      double* c = ck + (int64_t)(s / K) * cpw;
      for (int j = t; j < S.n; j += TPS) c[j] = S.xq[j];
      for (int d = t; d < S.ndyn; d += TPS) c[S.n + d] = S.ed[d];
    }
  }
  const double J = lin_block_sum<TPS>(jpart, red);
  if (t == 0) {
    Jout[inst] = J;
    status[inst] = flags[0];
  }
}

// Recomputation of every checkpoint segment at once, ahead of the backward
// pass: block (instance, segment) restores the segment's checkpoint and runs
// its K steps (the transient's LDS-resident step), writing each state
// (permuted order) into traj: [instance][nsteps + 1][n]. Segments are
// independent, so the whole trajectory costs about one segment's latency per
// wave of blocks instead of the transient's length; the backward pass then
// reads the states. Segment 0 also writes the state at step 0.
template <int TPS, int R>
__global__ void __launch_bounds__(TPS) k_adj_lin_rec(BatchDev B, int kind_slot, TranArgs ta, int K,
                                                     const double* ckpt, int64_t ck_stride, int cpw, double* traj,
                                                     int64_t traj_stride, const int32_t* status, double* srcv,
                                                     int nsrc_max) {
  extern __shared__ __align__(16) char smem[];
  const int t = threadIdx.x;
  const int nseg = (ta.nsteps + K - 1) / K;
  const int64_t slot = xcd_block();
  if (slot >= (int64_t)B.n_instances * nseg) return;
  const int inst = (int)(slot / nseg), sgi = (int)(slot % nseg);
  const PatDev& P = B.pats[B.inst_pat[kind_slot * B.n_instances + inst]];
  if (status[inst] != KG_OK) return;
  if (P.lin == nullptr || P.n > R * TPS || (srcv == nullptr && nsrc_max > 0)) return;  // the backward pass reports
  LinSys<TPS, R> S;
  const int words = reinterpret_cast<const LinHdr*>(P.lin)->words_fwd;
  int* flags = reinterpret_cast<int*>(S.setup(P, words, smem, ta, srcv, nsrc_max, inst));
  if (t == 0) flags[0] = KG_OK;
  __syncthreads();
  S.factorise(P, B.params + B.eoff[inst] * KG_NPARAM, B.ebase + B.eoff[inst], ta, &flags[0]);
  __syncthreads();
  if (flags[0] != KG_OK) return;
  S.cache();
  const int n = S.n;
  const int s0 = sgi * K, s1 = min(s0 + K, ta.nsteps);
  const double* c = ckpt + (int64_t)inst * ck_stride + (int64_t)sgi * cpw;
  double* tr = traj + (int64_t)inst * traj_stride;
  for (int j = t; j < n; j += TPS) {
    const double xv = c[j];
    S.xq[j] = xv;
    if (sgi == 0) tr[j] = xv;
  }
  for (int d = t; d < S.ndyn; d += TPS) S.ed[d] = c[n + d];
  S.prefetch(s0 + 1);
  __syncthreads();  // also publishes the source table within the block
  for (int s = s0 + 1; s <= s1; s++) {
    S.rhs(s);
    __syncthreads();
    S.solve();
    S.update();
    double* so = tr + (int64_t)s * n;
    for (int j = t; j < n; j += TPS) so[j] = S.xc[j];
    __syncthreads();
    S.swap();
  }
}

// Backward pass of the discrete adjoint from LDS (k_adj_bwd semantics): per
// checkpoint segment, last to first, recompute the segment's solutions into
// seg (permuted order), then run each step's adjoint backwards: the
// companions' adjoint loads, lambda = A^-T xbar by the transposed level
// solves, the gradient terms of every R, C and L and the companions' adjoint
// states. The recomputed state x_{s-1} is loaded into registers one step
// ahead and staged into LDS at the end of step s. xbar0 receives dJ/dx_0
// (krets numbering) for the operating-point term; grad: [instance][ne].
template <int TPS, int R>
__global__ void __launch_bounds__(TPS) k_adj_lin_bwd(BatchDev B, int kind_slot, TranArgs ta, int K,
                                                     const double* cvec, const double* ckpt, int64_t ck_stride,
                                                     int cpw, double* seg, int64_t seg_stride, double* grad,
                                                     int64_t grad_stride, double* xbar0, const int64_t* xop_off,
                                                     int32_t* status, double* srcv, int nsrc_max,
                                                     const double* xs_all, int64_t xs_inst_stride,
                                                     int64_t xs_step_stride) {
  extern __shared__ __align__(16) char smem[];
  const int t = threadIdx.x;
  const int inst = (int)xcd_block();
  if (inst >= B.n_instances) return;
  const PatDev& P = B.pats[B.inst_pat[kind_slot * B.n_instances + inst]];
  if (status[inst] != KG_OK) return;
  if (P.lin == nullptr || P.n > R * TPS || (srcv == nullptr && nsrc_max > 0)) {
    if (t == 0) status[inst] = KG_UNEXPECTED;
    return;
  }
  LinSys<TPS, R> S;
  const LinHdr* Hg = reinterpret_cast<const LinHdr*>(P.lin);
  char* q = S.setup(P, Hg->words, smem, ta, srcv, nsrc_max, inst);
  const LinHdr* H = reinterpret_cast<const LinHdr*>(smem);
  const int n = S.n, ndyn = S.ndyn, nadj = H->nadj, ne = H->ne, ntlev = H->ntlev, ntblev = H->ntblev;
  const uint16_t* tlvl_ptr = S.lw + H->o_tlvl_ptr;
  const ushort4* tfdesc = reinterpret_cast<const ushort4*>(S.lw + H->o_tfdesc);
  const uint16_t* tblvl_ptr = S.lw + H->o_tblvl_ptr;
  const ushort4* tbdesc = reinterpret_cast<const ushort4*>(S.lw + H->o_tbdesc);
  const uint16_t* tent = S.lw + H->o_tent;
  const uint16_t* adj = S.lw + H->o_adj;
  const int tl0 = H->tlev0_free ? 1 : 0;
  double* xbar = reinterpret_cast<double*>(q);  // adjoint load of the current step
  q += 8 * ((size_t)n + 1);
  double* xb0 = reinterpret_cast<double*>(q);  // its part that flows to the previous step
  q += 8 * ((size_t)n + 1);
  double* XA = reinterpret_cast<double*>(q);  // staged recomputed states
  q += 8 * ((size_t)n + 1);
  double* XB = reinterpret_cast<double*>(q);
  q += 8 * ((size_t)n + 1);
  double* sbar = reinterpret_cast<double*>(q);  // companion adjoint states
  q += 8 * ((size_t)ndyn + 1);
  double* gr = reinterpret_cast<double*>(q);  // gradient per element
  q += 8 * ((size_t)ne + 1);
  int* flags = reinterpret_cast<int*>(q + 8 * 32);
  if (t == 0) {
    flags[0] = KG_OK;
    XA[n] = XB[n] = 0.0;
  }
  __syncthreads();
  const double* prm = B.params + B.eoff[inst] * KG_NPARAM;
  S.factorise(P, prm, B.ebase + B.eoff[inst], ta, &flags[0]);
  __syncthreads();
  if (flags[0] != KG_OK) {
    if (t == 0) status[inst] = flags[0];
    return;
  }
  S.cache();
  // transposed-solve caches: position t + r*TPS of the U^T and L^T level
  // orders (level, row | first << 16, count | diagonal << 16, entries
  // column | slot << 16; padding points at the zero slots)
  int tf_lev[R], tb_lev[R];
  uint32_t tf_ip[R], tf_nd[R], tf_e[R][LIN_KE], tb_ip[R], tb_nd[R], tb_e[R][LIN_KE];
  double cv[R];
  {
    auto level_of = [&](const uint16_t* ptr, int nl, int g) {
      int l = 0;
      while (l + 1 < nl && (int)ptr[l + 1] <= g) l++;
      return l;
    };
    const uint32_t pad = (uint32_t)n | ((uint32_t)S.nnz << 16);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int g = t + r * TPS;
      const bool ok = g < n;
      tf_lev[r] = tb_lev[r] = -1;
      tf_ip[r] = tf_nd[r] = tb_ip[r] = tb_nd[r] = 0;
#pragma unroll
      for (int k = 0; k < LIN_KE; k++) tf_e[r][k] = tb_e[r][k] = pad;
      cv[r] = ok ? cvec[g] : 0.0;
      if (ok) {
        const ushort4 fd = tfdesc[g], bd = tbdesc[g];
        tf_lev[r] = level_of(tlvl_ptr, ntlev, g);
        tf_ip[r] = (uint32_t)fd.x | ((uint32_t)fd.y << 16);
        tf_nd[r] = (uint32_t)(fd.z - fd.y) | ((uint32_t)fd.w << 16);
        tb_lev[r] = level_of(tblvl_ptr, ntblev, g);
        tb_ip[r] = (uint32_t)bd.x | ((uint32_t)bd.y << 16);
        tb_nd[r] = (uint32_t)(bd.z - bd.y);
#pragma unroll
        for (int k = 0; k < LIN_KE; k++) {
          if (k < fd.z - fd.y)
            tf_e[r][k] = (uint32_t)tent[2 * (fd.y + k)] | ((uint32_t)tent[2 * (fd.y + k) + 1] << 16);
          if (k < bd.z - bd.y)
            tb_e[r][k] = (uint32_t)tent[2 * (bd.y + k)] | ((uint32_t)tent[2 * (bd.y + k) + 1] << 16);
        }
      }
    }
  }
  const double h = ta.h, al = ta.alpha;
  const bool trap = S.trap;
  // xbar_N = (h/2) c; no companion adjoint yet; gradients 0
#pragma unroll
  for (int r = 0; r < R; r++)
    if (t + r * TPS < n) {
      xbar[S.o_slot[r]] = 0.5 * h * cv[r];
      xb0[S.o_slot[r]] = 0.0;
    }
  for (int d = t; d <= ndyn; d += TPS) sbar[d] = 0.0;
  for (int e = t; e <= ne; e += TPS) gr[e] = 0.0;
  if (t == 0) xbar[n] = xb0[n] = 0.0;
  __syncthreads();
  const double* ck = ckpt + (int64_t)inst * ck_stride;
  double* sg = seg + (int64_t)inst * seg_stride;
  // every state given (xs_all: the checkpoints when K = 1, or the trajectory
  // k_adj_lin_rec recomputed): one segment over the whole run, whose states
  // are read where they lie, and nothing is recomputed
  const bool direct = xs_all != nullptr;
  const int KS = direct ? ta.nsteps : K;
  const double* sgp = direct ? xs_all + (int64_t)inst * xs_inst_stride : sg;  // the segment's states
  const int64_t sst = direct ? xs_step_stride : n;                            // and their stride
  const int nseg = (ta.nsteps + KS - 1) / KS;
  for (int sgi = nseg - 1; sgi >= 0; sgi--) {
    const int s0 = sgi * KS, s1 = min(s0 + KS, ta.nsteps);
    double* X1 = XA;
    double* Xs = XB;
    double* lam = S.xc;  // lambda lives in the recompute's free solution buffer
    if (direct) {
      for (int j = t; j < n; j += TPS) {
        X1[j] = sgp[(int64_t)s1 * sst + j];
        Xs[j] = sgp[(int64_t)(s1 - 1) * sst + j];
      }
      __syncthreads();
    } else {
      // restore the checkpoint and recompute the segment (solutions of steps
      // s0..s1 into seg, permuted order)
      {
        const double* c = ck + (int64_t)sgi * cpw;
        for (int j = t; j < n; j += TPS) {
          const double xv = c[j];
          S.xq[j] = xv;
          sg[j] = xv;
        }
        for (int d = t; d < ndyn; d += TPS) S.ed[d] = c[n + d];
      }
      S.prefetch(s0 + 1);
      __syncthreads();
      for (int s = s0 + 1; s <= s1; s++) {
        S.rhs(s);
        __syncthreads();
        S.solve();
        S.update();
        if (s < s1) {  // x_{s1} stays in LDS
          double* so = sg + (int64_t)(s - s0) * n;
          for (int j = t; j < n; j += TPS) so[j] = S.xc[j];
        }
        __syncthreads();
        S.swap();
      }
      // stage x_{s1} (LDS) and x_{s1-1} (the block's own stores: row 0 is the
      // checkpoint)
      for (int j = t; j < n; j += TPS) {
        X1[j] = S.xq[j];
        Xs[j] = sg[(int64_t)(s1 - 1 - s0) * n + j];
      }
      __syncthreads();
    }
    for (int s = s1 - 1; s >= s0; s--) {
      // x_{s-1} for the next step, into registers now
      double nx[R];
#pragma unroll
      for (int r = 0; r < R; r++)
        nx[r] = (s > s0 && t + r * TPS < n) ? sgp[(int64_t)(s - 1 - s0) * sst + t + r * TPS] : 0.0;
      // companions' adjoint loads of step s+1 (ic_{s+1} and vl_{s+1} depend
      // on x_{s+1}, x_s and ic_s)
      if (trap) {
        for (int d = t; d < ndyn; d += TPS) {
          const uint16_t* rd = S.dyn + d * LIN_DYN;
          const double sb1 = sbar[d];
          if (sb1 == 0.0) continue;
          const int p = (int)S.node(rd[1]), m = (int)S.node(rd[2]);
          if ((rd[0] >> 14) == 1) {
            const double a = S.cf[d] * sb1;  // 2C/h * sb1 (alpha = 2)
            if (p < n) { atomicAdd(&xbar[p], a); atomicAdd(&xb0[p], -a); }
            if (m < n) { atomicAdd(&xbar[m], -a); atomicAdd(&xb0[m], a); }
            const double d1 = X1[p] - X1[m], d0 = Xs[p] - Xs[m];
            gr[rd[0] & 0x3FFF] += sb1 * 2.0 / h * (d1 - d0);
          } else {
            if (p < n) atomicAdd(&xbar[p], sb1);
            if (m < n) atomicAdd(&xbar[m], -sb1);
          }
        }
        __syncthreads();
      }
      // lambda = A^-T xbar: U^T forward (level 0 fused with the copy), then
      // L^T backward (level 0 rows, when free of entries, are done)
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (tf_lev[r] < 0) continue;
        const int i = (int)(tf_ip[r] & 0xFFFF);
        lam[i] = tl0 && tf_lev[r] == 0 ? xbar[i] * S.v[tf_nd[r] >> 16] : xbar[i];
      }
      __syncthreads();
      for (int l = tl0; l < ntlev; l++) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (tf_lev[r] != l) continue;
          const int i = (int)(tf_ip[r] & 0xFFFF), q0 = (int)(tf_ip[r] >> 16), cnt = (int)(tf_nd[r] & 0xFFFF);
          double sum = lam[i];
          const double vd = S.v[tf_nd[r] >> 16];
          double vv[LIN_KE], xx[LIN_KE];
#pragma unroll
          for (int k = 0; k < LIN_KE; k++) {
            vv[k] = S.v[tf_e[r][k] >> 16];
            xx[k] = lam[tf_e[r][k] & 0xFFFF];
          }
#pragma unroll
          for (int k = 0; k < LIN_KE; k++) sum -= vv[k] * xx[k];
          for (int k = q0 + LIN_KE; k < q0 + cnt; k++) sum -= S.v[tent[2 * k + 1]] * lam[tent[2 * k]];
          lam[i] = sum * vd;
        }
        __syncthreads();
      }
      for (int l = tl0; l < ntblev; l++) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (tb_lev[r] != l) continue;
          const int i = (int)(tb_ip[r] & 0xFFFF), q0 = (int)(tb_ip[r] >> 16), cnt = (int)(tb_nd[r] & 0xFFFF);
          double sum = lam[i];
          double vv[LIN_KE], xx[LIN_KE];
#pragma unroll
          for (int k = 0; k < LIN_KE; k++) {
            vv[k] = S.v[tb_e[r][k] >> 16];
            xx[k] = lam[tb_e[r][k] & 0xFFFF];
          }
#pragma unroll
          for (int k = 0; k < LIN_KE; k++) sum -= vv[k] * xx[k];
          for (int k = q0 + LIN_KE; k < q0 + cnt; k++) sum -= S.v[tent[2 * k + 1]] * lam[tent[2 * k]];
          lam[i] = sum;
        }
        __syncthreads();
      }
      // element terms of step s -> s+1
      for (int a = t; a < nadj; a += TPS) {
        const uint16_t* ra = adj + a * LIN_ADJ;
        const int kind = ra[0] >> 14, e = ra[0] & 0x3FFF;
        const double v = prm[e * KG_NPARAM];
        const int p = (int)S.node(ra[1]), m = (int)S.node(ra[2]);
        if (kind == 2) {  // inductor: branch column c, branch row rc
          const int c = (int)ra[5], rc = (int)ra[6], d = (int)ra[7];
          const double lb = lam[rc];
          atomicAdd(&xb0[c], -al * (v / h) * lb);
          sbar[d] = trap ? -lb : 0.0;
          gr[e] += lb * (-al / h) * Xs[c] + al / h * lb * X1[c];
          continue;
        }
        const double el = lam[S.node(ra[3])] - lam[S.node(ra[4])];
        const double d1 = X1[p] - X1[m];
        if (kind == 0) {
          gr[e] += el * d1 / (v * v);
        } else {
          const int d = (int)ra[7];
          const double d0 = Xs[p] - Xs[m];
          const double aa = al * (v / h) * el;
          if (p < n) atomicAdd(&xb0[p], aa);
          if (m < n) atomicAdd(&xb0[m], -aa);
          sbar[d] = trap ? (-sbar[d] + el) : 0.0;
          gr[e] += el * al / h * d0 - al / h * el * d1;
        }
      }
      __syncthreads();
      // next step's adjoint load; stage x_{s-1}
      const double w = (s == 0) ? 0.5 * h : h;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (t + r * TPS < n) {
          const int j = S.o_slot[r];
          xbar[j] = xb0[j] + w * cv[r];
          xb0[j] = 0.0;
          X1[t + r * TPS] = nx[r];  // X1 is free now: it becomes x_{s-1}
        }
      __syncthreads();
      double* sw = X1;
      X1 = Xs;
      Xs = sw;
    }
  }
  // vl_0 = v(x_0): its adjoint flows into x_0
  if (trap) {
    for (int d = t; d < ndyn; d += TPS) {
      const uint16_t* rd = S.dyn + d * LIN_DYN;
      if ((rd[0] >> 14) != 2) continue;
      const double sb1 = sbar[d];
      const int p = (int)S.node(rd[1]), m = (int)S.node(rd[2]);
      if (p < n) atomicAdd(&xbar[p], sb1);
      if (m < n) atomicAdd(&xbar[m], -sb1);
    }
    __syncthreads();
  }
  double* xo = xbar0 + xop_off[inst];
#pragma unroll
  for (int r = 0; r < R; r++)
    if (t + r * TPS < n) xo[t + r * TPS] = xbar[S.o_slot[r]];
  double* g = grad + (int64_t)inst * grad_stride;
  for (int e = t; e < ne; e += TPS) g[e] = gr[e];
  if (t == 0) status[inst] = flags[0];
}

// Backward Euler with local-truncation-error step control (clean mode; no
// krets counterpart). Each system advances its own time grid; the LTE of
// every capacitor voltage and inductor current is estimated from the
// second divided difference of the last three accepted points,
// LTE = h^2/2 * |x''|, against reltol * max(|x_new|, |x_old|) + abstol.
// Rejected steps shrink h by max(0.25, 0.9/sqrt(ratio)); accepted steps
// grow it by up to 2x; a failed Newton solve divides h by 8.
// xop: clean operating points; tout: [instance][max_points];
// xout: krets numbering at xout_off[instance] + point * n.
// STAGED: every system of a block shares one plan (shared topology, or one
// system per block); it is copied into the first lay_al(4 * plan_ints) bytes
// of LDS once.
template <int TPS, bool STAGED>
__global__ void KG_LB(TPS) k_tran_lte(BatchDev B, int kind_slot, NewtonArgs na, LteArgs la, const double* xop,
                           const int64_t* xop_off, double* tout, double* xout, const int64_t* xout_off, int32_t* npts,
                           int32_t* newton_total, int32_t* rejected, int32_t* status, Lay lay, int plan_ints) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const int64_t slot0 = lay.sys_off + xcd_block() * spb;
  const int inst = (int)(slot0 + ls);
  const bool valid = inst < B.n_instances;
  const int ii = valid ? inst : 0;
  char* sys_lds = sys_region(smem, lay);
  const PatDev& Pg = B.pats[B.inst_pat[kind_slot * B.n_instances + ii]];
  PatDev Ps;
  if constexpr (STAGED) {
    const int first = slot0 < B.n_instances ? (int)slot0 : 0;
    const PatDev& G = B.pats[B.inst_pat[kind_slot * B.n_instances + first]];
    int32_t* pl = (int32_t*)smem;
    for (int i = threadIdx.x; i < G.total; i += blockDim.x) pl[i] = G.base[i];
    __syncthreads();
    Ps = rebase(G, pl);
    if (!lay.gws) sys_lds = smem + lay_al(sizeof(int32_t) * (size_t)plan_ints);
  }
  const PatDev& P = [&]() -> const PatDev& {
    if constexpr (STAGED) return Ps;
    else return Pg;
  }();
  size_t sb = sys_bytes(lay);
  double* red = (double*)(sys_lds + spb * sb);
  SysMem<double, double> M = carve<double, double>(sys_lds + ls * sb, lay);
  double* xc = M.xp;
  double* xpr = M.xp + M.xstride;
  const double* prm = B.params + B.eoff[ii] * KG_NPARAM;
  const double* eb = B.ebase + B.eoff[ii];
  const double* x0 = xop + xop_off[ii];
  if (t == 0) M.flags[0] = valid ? status[inst] : KG_OK;
  if (lay.stage_prm) {
    // every Newton iteration reads the element records: keep the system's in LDS
    const int np = P.ne * KG_NPARAM;
    for (int k = t; k < np; k += TPS) M.prm[k] = prm[k];
    for (int k = t; k < P.ne; k += TPS) M.prm[np + k] = eb[k];
    prm = M.prm;
    eb = M.prm + np;
  }
  __syncthreads();
  bool active = valid && M.flags[0] == KG_OK;
  int cnt = 0, nt = valid ? newton_total[inst] : 0, rej = 0;
  double tcur = 0.0, h = la.h0, tprev = 0.0;
  bool have_prev = false;
  if (active) {
    for (int j = t; j < P.n; j += TPS) {
      double v = x0[P.col_of[j]];
      xc[j] = v;
      xpr[j] = v;
    }
    double* o = xout + xout_off[ii];
    for (int j = t; j < P.n; j += TPS) o[j] = x0[j];
    if (t == 0) tout[(int64_t)ii * la.max_points] = 0.0;
    cnt = 1;
  }
  __syncthreads();
    StampArgs sa{AK_TRAN, 0.0, 1.0, 1.0, 0.0, -1, 0.0};
  const int64_t guard_max = 64 * (int64_t)la.max_points + 1000;
  for (int64_t guard = 0;; guard++) {
    if (!__syncthreads_or(active ? 1 : 0)) break;
    if (guard > guard_max) {
      if (active && t == 0) M.flags[0] = KG_MAX_ITERATIONS;
      break;
    }
    double tn = tcur;
    if (active) {
      h = fmin(h, la.hmax);
      if (tcur + h > la.tstop) h = la.tstop - tcur;
      tn = tcur + h;
    }
    sa.h = active ? h : 1.0;
    sa.tsrc = tn;
    // Newton solve of the backward-Euler step from xc
    if (active) {
      for (int i = t; i < P.n; i += TPS) M.g[i] = xc[i];
      for (int e = t; e < P.ne; e += TPS)
        if (P.etype[e] == KG_D) {
          int p = P.ep[e], m = P.em[e];
          M.ed[e * ESTRIDE + 2] = (p >= 0 ? xc[p] : 0.0) - (m >= 0 ? xc[m] : 0.0);
        }
    }
    __syncthreads();
    bool done = !active;
    int its = 0;
    double anp = 0.0;  // this thread's part of ||A||_inf of the last factorisation
    for (int it = 0; it < na.maxiter; it++) {
      bool a2 = !done;
      if (t == 0) M.flags[4] = 0;
      __syncthreads();
      eval_diodes<TPS>(P, prm, M.g, M.ed, false, false, &M.flags[4], t, a2, &sa, xc);
      __syncthreads();
      stamp_matrix<double, double, TPS>(P, M.v, prm, eb, sa, M.ed, t, a2);
      stamp_rhs<double, TPS>(P, M.b, prm, sa, M.ed, xc, t, a2);
      __syncthreads();
      if (a2) anp = 0.0;  // a system that converged earlier keeps its last factorisation's
      factor<double, TPS, true>(P, M.v, t, a2, &M.flags[0], &anp);
      solve_refine<double, double, TPS>(P, M.v, M.b, M.x, M.w, prm, eb, sa, M.ed, 0, red, ls, t, a2);
      __syncthreads();
      bool stop;
      if (!P.nonlinear) {
        stop = true;
        sys_max<TPS>(0.0, red, ls, t);
      } else {
        bool cv = newton_converged<TPS>(P, M.x, M.g, na.reltol, na.iabstol, na.vabstol, red, ls, t, a2);
        stop = cv && !M.flags[4];
      }
      if (M.flags[0] != KG_OK) stop = true;
      if (a2 && !stop)
        for (int i = t; i < P.n; i += TPS) M.g[i] = M.x[i];
      __syncthreads();
      if (a2 && stop) {
        done = true;
        its = it + 1;
      }
      if (__syncthreads_and(done ? 1 : 0)) break;
    }
    bool ok = active && done && M.flags[0] == KG_OK;
    {  // static-pivot guard on the factors of the step's solution (b: scratch)
      const double bnd = pivot_guard<TPS>(P, M.v, anp, M.b, red, ls, t, ok);
      if (ok && !(bnd <= KG_GUARD_STATIC)) {
        if (t == 0) M.flags[0] = KG_DECOMPOSITION;
        ok = false;
      }
      __syncthreads();
    }
    if (active && M.flags[0] != KG_OK) active = false;
    if (active && !done) {  // Newton failure: retry with h / 8
      rej++;
      h /= 8.0;
      if (h < la.hmin) {
        if (t == 0) M.flags[0] = KG_MAX_ITERATIONS;
        active = false;
      }
    }
    if (ok) nt += its;
    // LTE ratio over capacitor voltages and inductor currents
    double ratio = 0.0;
    if (ok && have_prev) {
      double hp = tcur - tprev;
      for (int e = t; e < P.ne; e += TPS) {
        int ty = P.etype[e];
        double a, b0, bm;
        if (ty == KG_C || (ty == KG_D && has_jq(prm + e * KG_NPARAM))) {  // junction charge: its voltage
          int p = P.ep[e], m = P.em[e];
          a = (p >= 0 ? M.x[p] : 0.0) - (m >= 0 ? M.x[m] : 0.0);
          b0 = (p >= 0 ? xc[p] : 0.0) - (m >= 0 ? xc[m] : 0.0);
          bm = (p >= 0 ? xpr[p] : 0.0) - (m >= 0 ? xpr[m] : 0.0);
        } else if (ty == KG_L) {
          int c = P.ec[e];
          a = M.x[c];
          b0 = xc[c];
          bm = xpr[c];
        } else {
          continue;
        }
        double dd2 = 2.0 * ((a - b0) / h - (b0 - bm) / hp) / (h + hp);
        double lte = 0.5 * h * h * fabs(dd2);
        double tol = la.lte_reltol * fmax(fabs(a), fabs(b0)) + la.lte_abstol;
        ratio = fmax(ratio, lte / tol);
      }
    }
    ratio = sys_max<TPS>(ratio, red, ls, t);
    bool accept = false;
    double hnext = h;
    if (ok) {
      accept = true;
      if (have_prev) {
        if (ratio > 1.0) {
          rej++;
          h *= fmax(0.25, 0.9 / sqrt(ratio));
          accept = false;
          if (h < la.hmin) {
            if (t == 0) M.flags[0] = KG_MAX_ITERATIONS;
            active = false;
          }
        } else {
          hnext = ratio > 0.0 ? h * fmin(2.0, fmax(0.25, 0.9 / sqrt(ratio))) : 2.0 * h;
        }
      }
      if (accept && cnt >= la.max_points) {
        if (t == 0) M.flags[0] = KG_UNEXPECTED;
        active = false;
        accept = false;
      }
    }
    __syncthreads();
    if (accept) {
      for (int i = t; i < P.n; i += TPS) {
        xpr[i] = xc[i];
        xc[i] = M.x[i];
      }
      double* o = xout + xout_off[ii] + (int64_t)cnt * P.n;
      for (int j = t; j < P.n; j += TPS) o[P.col_of[j]] = M.x[j];
      if (t == 0) tout[(int64_t)ii * la.max_points + cnt] = tn;
      cnt++;
      tprev = tcur;
      tcur = tn;
      h = hnext;
      have_prev = true;
      if (!(tcur < la.tstop * (1.0 - 1e-12))) active = false;
    }
    __syncthreads();
  }
  if (valid && t == 0) {
    status[inst] = M.flags[0];
    npts[inst] = cnt;
    newton_total[inst] = nt;
    rejected[inst] = rej;
  }
}



// Re-points every array of a plan into a copy of its buffer at `lds`.

// Persistent AC sweep for batches with one shared topology: a grid of a few
// blocks per CU, each walking a contiguous chunk of work items. A work item is
// SPB consecutive frequencies of one instance. The plan is copied into LDS once
// per block and the instance's element records whenever the instance changes,
// so the gather and the factor program read LDS instead of issuing dependent
// global loads.
// LDS: [plan (plan_ints, 16-aligned)] [element records ne*KG_NPARAM + ebase ne]
//      [SPB systems] [reduction scratch].
template <class F, class A, int TPS>
__global__ void KG_LB(TPS) k_ac_p(BatchDev B, int kind_slot, const double* freqs, int nf, const double* xop,
                       const int64_t* xop_off, double* out, const int64_t* out_off, int32_t* status, double* err,
                       int refine, Lay lay, int plan_ints, int64_t chunk, unsigned long long* fixq,
                       int64_t fixcap) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const PatDev& G = B.pats[B.inst_pat[kind_slot * B.n_instances]];
  int32_t* pl = (int32_t*)smem;
  size_t off = lay_al(sizeof(int32_t) * (size_t)plan_ints);
  double* iprm = (double*)(smem + off);
  off += lay_al(sizeof(double) * (size_t)G.ne * KG_NPARAM);
  double* ieb = (double*)(smem + off);
  off += lay_al(sizeof(double) * (size_t)G.ne);
  size_t sb = sys_bytes(lay);
  char* sysbase = smem + off;
  double* red = (double*)(sysbase + spb * sb);
  SysMem<F, A> M = carve<F, A>(sysbase + ls * sb, lay);
  for (int i = threadIdx.x; i < G.total; i += blockDim.x) pl[i] = G.base[i];
  __syncthreads();
  const PatDev P = rebase(G, pl);
  const int64_t gpi = (nf + spb - 1) / spb;  // work items per instance
  const int64_t nitems = gpi * B.n_instances;
  const int64_t i0 = xcd_block() * chunk, i1 = min(i0 + chunk, nitems);
  int cur = -1;
  for (int64_t item = i0; item < i1; item++) {
    const int inst = (int)(item / gpi);
    const int fi = (int)((item % gpi) * spb + ls);
    const bool valid = fi < nf;
    if (inst != cur) {  // stage the instance's element records
      __syncthreads();
      const double* gp = B.params + B.eoff[inst] * KG_NPARAM;
      const double* ge = B.ebase + B.eoff[inst];
      for (int i = threadIdx.x; i < G.ne * KG_NPARAM; i += blockDim.x) iprm[i] = gp[i];
      for (int i = threadIdx.x; i < G.ne; i += blockDim.x) ieb[i] = ge[i];
      cur = inst;
      __syncthreads();
    }
    if (t == 0) M.flags[0] = KG_OK;
    if (valid && P.has_diode)
      for (int e = t; e < P.ne; e += TPS)
        if (P.etype[e] == KG_D) {
          const double* x0 = xop + xop_off[inst];
          int p = P.ep[e], m = P.em[e];
          double vp = p >= 0 ? x0[P.col_of[p]] : 0.5, vm = m >= 0 ? x0[P.col_of[m]] : 0.0;
          const double* pe = iprm + e * KG_NPARAM;
          DiodeLin d = diode_eval(vp - vm, pe[0], pe[1], true, 0.0);
          M.ed[e * ESTRIDE] = d.g;
          M.ed[e * ESTRIDE + 1] = has_jq(pe) ? junction_charge(vp - vm, pe, d.ieq + d.g * (vp - vm), d.g).c : 0.0;
        }
    __syncthreads();
    StampArgs sa{AK_AC, valid ? freqs[fi] : 1.0, 1.0, 1.0, 0.0, -1, 0.0};
    stamp_matrix<F, A, TPS>(P, M.v, iprm, ieb, sa, M.ed, t, valid);
    stamp_rhs<A, TPS>(P, M.b, iprm, sa, M.ed, nullptr, t, valid);
    __syncthreads();
    factor<F, TPS>(P, M.v, t, valid, &M.flags[0]);
    double est =
        solve_refine<F, A, TPS>(P, M.v, M.b, M.x, M.w, iprm, ieb, sa, M.ed, refine, red, ls, t, valid, M.r);
    __syncthreads();
    constexpr bool mixed = !std::is_same<F, A>::value;
    if constexpr (!mixed) {
      est = guard_omega<A, TPS>(P, M.b, M.x, iprm, ieb, sa, M.ed, red, ls, t, valid);
    } else if (!(guard_eta<A, TPS>(P, M.b, M.x, iprm, ieb, sa, M.ed, red, ls, t, valid) <= KG_GUARD_MIXED_ETA)) {
      est = INFINITY;  // the returned solution misses the normwise guard: to the fixup
    }
    if (valid) {
      double* o = out + out_off[inst] + (int64_t)fi * P.n * 2;
      for (int j = t; j < P.n; j += TPS) {
        int c = P.col_of[j];
        o[2 * c] = re_of(M.x[j]);
        o[2 * c + 1] = im_of(M.x[j]);
      }
      if (t == 0) ac_finish(M.flags[0], est, mixed, fixq, fixcap, (int64_t)inst * nf + fi, status, err, inst);
    }
  }
}

// ------------------------------------------------------------- AC fixup
// Solves the AC systems queued by the sweep kernels' accuracy guard again,
// the way krets' sparse LU does it: with row partial pivoting chosen at the
// point itself (crates/krets-solver/src/solver/ac.rs:67-95 calls faer's
// sp_lu per frequency). The system (at most KG_FIX_NMAX unknowns) is stamped
// into a dense complex128 matrix in LDS in the plan's permuted numbering and
// factored by one block of KG_FIX_NT threads: per column a one-wave pivot
// search, a row swap, the multipliers and the rank-1 update of the trailing
// matrix. Up to two FP64 refinement steps follow, with residuals from the
// sparse stamps. Persistent over the queue: block k takes queued systems k,
// k + gridDim.x, ..., so the grid never depends on the count.
// err_mode 0 (FP64 calls) publishes the final backward error, 1 (mixed) the
// size of the last refinement correction relative to x.
constexpr int KG_FIX_NT = 256;

__device__ __forceinline__ double wave_max_idx(double v, int& idx) {
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const int oi = __shfl_xor(idx, off);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  return v;
}

__device__ __forceinline__ c128 wave_sum_c(c128 v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.re += __shfl_xor(v.re, off);
    v.im += __shfl_xor(v.im, off);
  }
  return v;
}

__device__ double block_max(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = KG_FIX_NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = fmax(red[t], red[t + s]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Forward and back substitution with the pivoted dense LU (unit lower L
// below the diagonal, U on and above it) on w, in place; one wave works, the
// block synchronises around it.
__device__ void fix_lu_solve(const c128* LU, int ld, int n, const int32_t* piv, c128* w) {
  const int t = threadIdx.x;
  if (t < 64) {
    for (int k = 0; k < n; k++) {  // the row swaps of the factorisation, in order
      const int p = piv[k];
      if (t == 0 && p != k) {
        const c128 tmp = w[k];
        w[k] = w[p];
        w[p] = tmp;
      }
    }
    for (int i = 0; i < n; i++) {
      c128 acc{0.0, 0.0};
      for (int j = t; j < i; j += 64) acc += LU[i * ld + j] * w[j];
      acc = wave_sum_c(acc);
      if (t == 0) w[i] = w[i] - acc;
    }
    for (int i = n - 1; i >= 0; i--) {
      c128 acc{0.0, 0.0};
      for (int j = i + 1 + t; j < n; j += 64) acc += LU[i * ld + j] * w[j];
      acc = wave_sum_c(acc);
      if (t == 0) w[i] = (w[i] - acc) * rcp(LU[i * ld + i]);
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(KG_FIX_NT) k_ac_fix(BatchDev B, int kind_slot, const double* freqs, int nf,
                                                      const double* xop, const int64_t* xop_off, double* out,
                                                      const int64_t* out_off, int32_t* status, double* err,
                                                      int err_mode, const unsigned long long* fixq, int64_t fixcap,
                                                      int nmax, int ne_max, int nskip, int defer, char* gws,
                                                      int64_t gstride) {
  // Re-solves the queued systems of nskip < n <= nmax unknowns. The first
  // launch works in LDS (nskip = 0); with defer, larger systems are left to a
  // second launch that works from a global workspace region per block
  // (gws + block * gstride, the same carve-up) with nskip the first one's
  // nmax; a system beyond every launch's nmax reports KG_DECOMPOSITION.
  extern __shared__ __align__(16) char smem[];
  const int t = threadIdx.x;
  char* base = gws ? gws + (int64_t)blockIdx.x * gstride : smem;
  c128* LU = (c128*)base;                      // [nmax][nmax]
  c128* bb = LU + (size_t)nmax * nmax;         // right-hand side (plan row order)
  c128* x = bb + nmax;                         // solution (plan column order)
  c128* w = x + nmax;                          // work vector
  double* red = (double*)(w + nmax);           // [KG_FIX_NT]
  double* ed = red + KG_FIX_NT;                // [ESTRIDE * ne_max] diode small-signal state
  int32_t* piv = (int32_t*)(ed + (size_t)ESTRIDE * ne_max);  // [nmax]
  int32_t* sh = piv + nmax;                    // [0] pivot row, [1] status
  const int64_t cnt = min((int64_t)fixq[0], fixcap);
  for (int64_t q = blockIdx.x; q < cnt; q += gridDim.x) {
    const int64_t sys = (int64_t)fixq[1 + q];
    const int inst = (int)(sys / nf), fi = (int)(sys % nf);
    const PatDev& P = B.pats[B.inst_pat[kind_slot * B.n_instances + inst]];
    const int n = P.n;
    if (n <= nskip) continue;  // re-solved by the earlier launch
    if (n > nmax || P.ne > ne_max) {  // block-uniform
      if (!defer && t == 0) atomicMax(&status[inst], KG_DECOMPOSITION);
      continue;
    }
    const double* prm = B.params + B.eoff[inst] * KG_NPARAM;
    const double* eb = B.ebase + B.eoff[inst];
    for (int e = t; e < P.ne; e += KG_FIX_NT)
      if (P.etype[e] == KG_D) {  // small-signal conductance at the operating point (stampable.rs:614-641)
        const double* x0 = xop + xop_off[inst];
        const int p = P.ep[e], m = P.em[e];
        const double vp = p >= 0 ? x0[P.col_of[p]] : 0.5, vm = m >= 0 ? x0[P.col_of[m]] : 0.0;
        const double* pe = prm + e * KG_NPARAM;
        const DiodeLin d = diode_eval(vp - vm, pe[0], pe[1], true, 0.0);
        ed[e * ESTRIDE] = d.g;
        ed[e * ESTRIDE + 1] = has_jq(pe) ? junction_charge(vp - vm, pe, d.ieq + d.g * (vp - vm), d.g).c : 0.0;
      }
    for (int k = t; k < n * nmax; k += KG_FIX_NT) LU[k] = c128{0.0, 0.0};
    if (t == 0) sh[1] = KG_OK;
    __syncthreads();
    const StampArgs sa{AK_AC, freqs[fi], 1.0, 1.0, 0.0, -1, 0.0};
    for (int i = t; i < n; i += KG_FIX_NT) {
      for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++) {
        c128 a{0.0, 0.0};
        for (int c = P.scon_ptr[s]; c < P.scon_ptr[s + 1]; c++) a += mat_contrib<c128>(P.scon[c], P, prm, eb, sa, ed);
        LU[i * nmax + P.col[s]] = a;
      }
      c128 r{0.0, 0.0};
      for (int c = P.rcon_ptr[i]; c < P.rcon_ptr[i + 1]; c++) r += rhs_contrib<c128>(P.rcon[c], P, prm, sa, ed, nullptr);
      bb[i] = r;
    }
    __syncthreads();
    // LU with row partial pivoting
    for (int k = 0; k < n; k++) {
      if (t < 64) {
        double best = -1.0;
        int bi = k;
        for (int i = k + t; i < n; i += 64) {
          const double m = c1(LU[i * nmax + k]);
          if (m > best) {
            best = m;
            bi = i;
          }
        }
        best = wave_max_idx(best, bi);
        if (t == 0) {
          piv[k] = bi;
          sh[0] = bi;
          if (!(best > 0.0) || !isfinite(best)) sh[1] = KG_DECOMPOSITION;
        }
      }
      __syncthreads();
      const int p = sh[0];
      if (p != k)
        for (int j = t; j < n; j += KG_FIX_NT) {
          const c128 tmp = LU[k * nmax + j];
          LU[k * nmax + j] = LU[p * nmax + j];
          LU[p * nmax + j] = tmp;
        }
      __syncthreads();
      const c128 rk = rcp(LU[k * nmax + k]);
      for (int i = k + 1 + t; i < n; i += KG_FIX_NT) LU[i * nmax + k] = LU[i * nmax + k] * rk;
      __syncthreads();
      const int m = n - k - 1;
      for (int idx = t; idx < m * m; idx += KG_FIX_NT) {
        const int i = k + 1 + idx / m, j = k + 1 + idx % m;
        LU[i * nmax + j] -= LU[i * nmax + k] * LU[k * nmax + j];
      }
      __syncthreads();
    }
    const bool ok = sh[1] == KG_OK;
    for (int i = t; i < n; i += KG_FIX_NT) {
      w[i] = bb[i];
      x[i] = c128{0.0, 0.0};
    }
    __syncthreads();
    if (ok) fix_lu_solve(LU, nmax, n, piv, w);
    // x = w; then refinement steps with FP64 residuals from the sparse stamps
    double eta = 0.0, corr = 0.0;
    for (int it = 0; it < 3; it++) {
      if (ok)
        for (int i = t; i < n; i += KG_FIX_NT) x[i] += w[i];
      __syncthreads();
      double rm = 0.0, am = 0.0, bm = 0.0, xm = 0.0, dm = 0.0;
      for (int i = t; i < n; i += KG_FIX_NT) {
        c128 r = bb[i];
        double ar = 0.0;
        for (int s = P.row_ptr[i]; s < P.row_ptr[i + 1]; s++) {
          const int q0 = P.scon_ptr[s], q1 = P.scon_ptr[s + 1];
          if (q0 == q1) continue;
          c128 a{0.0, 0.0};
          for (int c = q0; c < q1; c++) a += mat_contrib<c128>(P.scon[c], P, prm, eb, sa, ed);
          r -= a * x[P.col[s]];
          ar += c1(a);
        }
        rm = fmax(rm, c1(r));
        am = fmax(am, ar);
        bm = fmax(bm, c1(bb[i]));
        xm = fmax(xm, c1(x[i]));
        dm = fmax(dm, c1(w[i]));
        w[i] = r;  // next correction's right-hand side (its own row: no other thread reads it)
      }
      rm = block_max(rm, red);
      am = block_max(am, red);
      bm = block_max(bm, red);
      xm = block_max(xm, red);
      dm = block_max(dm, red);
      const double den = am * xm + bm;
      eta = den > 0.0 ? rm / den : (rm > 0.0 ? 1.0 : 0.0);
      corr = xm > 0.0 ? dm / xm : 0.0;
      // stop once eta is at rounding level; the mixed figure needs one
      // refinement step (its correction estimates the first solve's error)
      if (!ok || it == 2 || (eta <= 4.0 * 2.220446049250313e-16 && (err_mode == 0 || it >= 1))) break;
      fix_lu_solve(LU, nmax, n, piv, w);
    }
    if (t == 0) {
      int32_t st = sh[1];
      if (st == KG_OK && !(eta <= KG_FIX_FAIL)) st = KG_DECOMPOSITION;
      if (st != KG_OK) atomicMax(&status[inst], st);
      const double q2 = err_mode ? corr : eta;
      if (err && q2 > 0.0)
        atomicMax(reinterpret_cast<unsigned long long*>(&err[inst]), (unsigned long long)__double_as_longlong(q2));
    }
    double* o = out + out_off[inst] + (int64_t)fi * n * 2;
    for (int j = t; j < n; j += KG_FIX_NT) {
      const int c = P.col_of[j];
      o[2 * c] = x[j].re;
      o[2 * c + 1] = x[j].im;
    }
    __syncthreads();
  }
}

// LDS bytes of k_ac_fix for systems of up to nmax unknowns and ne_max elements.
size_t ac_fix_lds(int nmax, int ne_max) {
  return sizeof(c128) * ((size_t)nmax * nmax + 3 * (size_t)nmax) + 8 * (size_t)KG_FIX_NT +
         8 * (size_t)ESTRIDE * ne_max + 4 * ((size_t)nmax + 2);
}

// ------------------------------------------------------------- adjoint
template <int TPS>
__device__ double sys_sum(double v, double* red, int ls, int t) {
  if constexpr (TPS == 1) {
    return v;
  } else {
    red[ls * TPS + t] = v;
    __syncthreads();
    for (int s = TPS / 2; s > 0; s >>= 1) {
      if (t < s) red[ls * TPS + t] += red[ls * TPS + t + s];
      __syncthreads();
    }
    double r = red[ls * TPS];
    __syncthreads();
    return r;
  }
}

// Forward pass of the discrete adjoint (linear circuits, clean BE or
// trapezoidal): the transient of k_tran plus the objective
// J = sum_s w_s c.x_s (trapezoid weights) and a checkpoint of the state
// (solution and companion states) every K steps.
// ckpt: [instance][ncp][cpw] with cpw >= n + ne; cvec in krets numbering.
template <class F, class A, int TPS, bool STAGED>
__global__ void KG_LB(TPS) k_adj_fwd(BatchDev B, int kind_slot, AdjArgs aa, const double* cvec, const double* xop,
                          const int64_t* xop_off, double* ckpt, int64_t ck_stride, int cpw, double* Jout,
                          int32_t* status, Lay lay, int plan_ints) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const int inst = (int)(xcd_block() * spb + ls);
  const bool valid = inst < B.n_instances;
  const int ii = valid ? inst : 0;
  char* sys_lds = smem;
  const PatDev& Pg = B.pats[B.inst_pat[kind_slot * B.n_instances + ii]];
  PatDev Ps;
  if constexpr (STAGED) {  // shared topology: one plan for the whole grid
    const PatDev& G = B.pats[B.inst_pat[kind_slot * B.n_instances]];
    int32_t* pl = (int32_t*)smem;
    for (int i = threadIdx.x; i < G.total; i += blockDim.x) pl[i] = G.base[i];
    __syncthreads();
    Ps = rebase(G, pl);
    sys_lds = smem + lay_al(sizeof(int32_t) * (size_t)plan_ints);
  }
  const PatDev& P = [&]() -> const PatDev& {
    if constexpr (STAGED) return Ps;
    else return Pg;
  }();
  size_t sb = sys_bytes(lay);
  double* red = (double*)(sys_lds + spb * sb);
  SysMem<F, A> M = carve<F, A>(sys_lds + ls * sb, lay);
  const double* prm = B.params + B.eoff[ii] * KG_NPARAM;
  const double* eb = B.ebase + B.eoff[ii];
  const double* x0 = xop + xop_off[ii];
  double* ck = ckpt + (int64_t)ii * ck_stride;
  if (t == 0) M.flags[0] = valid ? status[inst] : KG_OK;
  __syncthreads();
  bool act = valid && M.flags[0] == KG_OK;
  const bool trap = aa.trap != 0;
  StampArgs sa{AK_TRAN, 0.0, aa.h, aa.alpha, 0.0, -1, 0.0};
  if (act) {
    for (int j = t; j < P.n; j += TPS) M.xp[j] = x0[P.col_of[j]];
    for (int e = t; e < P.ne; e += TPS) {
      int p = P.ep[e], m = P.em[e];
      double vd = (p >= 0 ? x0[P.col_of[p]] : 0.0) - (m >= 0 ? x0[P.col_of[m]] : 0.0);
      M.ed[e * ESTRIDE] = (trap && P.etype[e] == KG_L) ? vd : 0.0;
    }
  }
  __syncthreads();
  stamp_matrix<F, A, TPS>(P, M.v, prm, eb, sa, M.ed, t, act);
  __syncthreads();
  factor<F, TPS>(P, M.v, t, act, &M.flags[0]);
  __syncthreads();
  double jpart = 0.0;
  for (int s = 0; s <= aa.nsteps; s++) {
    act = valid && M.flags[0] == KG_OK;
    if (s > 0) {
      sa.tsrc = (double)s * aa.h;
      stamp_rhs<A, TPS>(P, M.b, prm, sa, M.ed, M.xp, t, act);
      __syncthreads();
      solve_refine<F, A, TPS>(P, M.v, M.b, M.x, M.w, prm, eb, sa, M.ed, aa.refine, red, ls, t, act, M.r);
      __syncthreads();
      if (act && trap)
        for (int e = t; e < P.ne; e += TPS) {
          int ty = P.etype[e];
          if (ty != KG_C && ty != KG_L) continue;
          int p = P.ep[e], m = P.em[e];
          double vn = (p >= 0 ? re_of(M.x[p]) : 0.0) - (m >= 0 ? re_of(M.x[m]) : 0.0);
          if (ty == KG_C) {
            double vo = (p >= 0 ? M.xp[p] : 0.0) - (m >= 0 ? M.xp[m] : 0.0);
            M.ed[e * ESTRIDE] = 2.0 * prm[e * KG_NPARAM] / aa.h * (vn - vo) - M.ed[e * ESTRIDE];
          } else {
            M.ed[e * ESTRIDE] = vn;
          }
        }
      __syncthreads();
      if (act)
        for (int j = t; j < P.n; j += TPS) M.xp[j] = re_of(M.x[j]);
      __syncthreads();
    }
    double w = (s == 0 || s == aa.nsteps) ? 0.5 * aa.h : aa.h;
    if (act)
      for (int j = t; j < P.n; j += TPS) jpart += w * cvec[P.col_of[j]] * M.xp[j];
    if (act && s % aa.K == 0) {
      double* c = ck + (int64_t)(s / aa.K) * cpw;
      for (int j = t; j < P.n; j += TPS) c[j] = M.xp[j];
      for (int e = t; e < P.ne; e += TPS) c[P.n + e] = M.ed[e * ESTRIDE];
    }
  }
  double J = sys_sum<TPS>(jpart, red, ls, t);
  if (valid && t == 0) {
    Jout[inst] = J;
    status[inst] = M.flags[0];
  }
}

// Backward pass: per checkpoint segment (last to first) recompute the
// segment's solutions, then run the discrete adjoint of each step backwards:
// lambda = A^-T xbar with the stored LU (transposed solves), gradient terms of
// every R, C and L, and the adjoint states of the trapezoidal companions.
// xbar0 receives dJ/dx_0 (krets numbering) for the operating-point adjoint.
template <class F, class A, int TPS, bool STAGED>
__global__ void KG_LB(TPS) k_adj_bwd(BatchDev B, int kind_slot, AdjArgs aa, const double* cvec, const double* ckpt,
                          int64_t ck_stride, int cpw, double* seg, int64_t seg_stride, double* grad,
                          int64_t grad_stride, double* xbar0, const int64_t* xop_off, int32_t* status, Lay lay,
                          int plan_ints) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const int inst = (int)(xcd_block() * spb + ls);
  const bool valid = inst < B.n_instances;
  const int ii = valid ? inst : 0;
  char* sys_lds = smem;
  const PatDev& Pg = B.pats[B.inst_pat[kind_slot * B.n_instances + ii]];
  PatDev Ps;
  if constexpr (STAGED) {  // shared topology: one plan for the whole grid
    const PatDev& G = B.pats[B.inst_pat[kind_slot * B.n_instances]];
    int32_t* pl = (int32_t*)smem;
    for (int i = threadIdx.x; i < G.total; i += blockDim.x) pl[i] = G.base[i];
    __syncthreads();
    Ps = rebase(G, pl);
    sys_lds = smem + lay_al(sizeof(int32_t) * (size_t)plan_ints);
  }
  const PatDev& P = [&]() -> const PatDev& {
    if constexpr (STAGED) return Ps;
    else return Pg;
  }();
  size_t sb = sys_bytes(lay);
  double* red = (double*)(sys_lds + spb * sb);
  SysMem<F, A> M = carve<F, A>(sys_lds + ls * sb, lay);
  double* xbar = M.xp + M.xstride;
  double* xb0 = M.xp + 2 * M.xstride;
  double* lam = M.xp + 3 * M.xstride;
  double* rr = M.xp + 4 * M.xstride;
  const double* prm = B.params + B.eoff[ii] * KG_NPARAM;
  const double* eb = B.ebase + B.eoff[ii];
  const double* ck = ckpt + (int64_t)ii * ck_stride;
  double* sg = seg + (int64_t)ii * seg_stride;
  if (t == 0) M.flags[0] = valid ? status[inst] : KG_OK;
  __syncthreads();
  bool act = valid && M.flags[0] == KG_OK;
  const bool trap = aa.trap != 0;
  const double al = aa.alpha, h = aa.h;
  StampArgs sa{AK_TRAN, 0.0, h, al, 0.0, -1, 0.0};
  stamp_matrix<F, A, TPS>(P, M.v, prm, eb, sa, M.ed, t, act);
  __syncthreads();
  factor<F, TPS>(P, M.v, t, act, &M.flags[0]);
  if (act) {
    for (int j = t; j < P.n; j += TPS) {
      xbar[j] = 0.5 * h * cvec[P.col_of[j]];
      xb0[j] = 0.0;
    }
    for (int e = t; e < P.ne; e += TPS) {
      M.ed[e * ESTRIDE + 1] = 0.0;
      M.ed[e * ESTRIDE + 2] = 0.0;
    }
  }
  __syncthreads();
  const int nseg = (aa.nsteps + aa.K - 1) / aa.K;
  for (int sgi = nseg - 1; sgi >= 0; sgi--) {
    const int s0 = sgi * aa.K, s1 = min(s0 + aa.K, aa.nsteps);
    // restore the checkpoint and recompute the segment
    if (act) {
      const double* c = ck + (int64_t)sgi * cpw;
      for (int j = t; j < P.n; j += TPS) {
        M.xp[j] = c[j];
        sg[j] = c[j];
      }
      for (int e = t; e < P.ne; e += TPS) M.ed[e * ESTRIDE] = c[P.n + e];
    }
    __syncthreads();
    for (int s = s0 + 1; s <= s1; s++) {
      sa.tsrc = (double)s * h;
      stamp_rhs<A, TPS>(P, M.b, prm, sa, M.ed, M.xp, t, act);
      __syncthreads();
      solve_refine<F, A, TPS>(P, M.v, M.b, M.x, M.w, prm, eb, sa, M.ed, aa.refine, red, ls, t, act, M.r);
      __syncthreads();
      if (act && trap)
        for (int e = t; e < P.ne; e += TPS) {
          int ty = P.etype[e];
          if (ty != KG_C && ty != KG_L) continue;
          int p = P.ep[e], m = P.em[e];
          double vn = (p >= 0 ? re_of(M.x[p]) : 0.0) - (m >= 0 ? re_of(M.x[m]) : 0.0);
          if (ty == KG_C) {
            double vo = (p >= 0 ? M.xp[p] : 0.0) - (m >= 0 ? M.xp[m] : 0.0);
            M.ed[e * ESTRIDE] = 2.0 * prm[e * KG_NPARAM] / h * (vn - vo) - M.ed[e * ESTRIDE];
          } else {
            M.ed[e * ESTRIDE] = vn;
          }
        }
      __syncthreads();
      if (act)
        for (int j = t; j < P.n; j += TPS) {
          double v = re_of(M.x[j]);
          M.xp[j] = v;
          sg[(int64_t)(s - s0) * P.n + j] = v;
        }
      __syncthreads();
    }
    // adjoint steps s1-1 .. s0
    for (int s = s1 - 1; s >= s0; s--) {
      const double* x1 = sg + (int64_t)(s + 1 - s0) * P.n;
      const double* xo = sg + (int64_t)(s - s0) * P.n;
      if (act && trap)  // ic_{s+1} and vl_{s+1} depend on x_{s+1} (and x_s, ic_s)
        for (int e = t; e < P.ne; e += TPS) {
          int ty = P.etype[e], p = P.ep[e], m = P.em[e];
          double sb1 = M.ed[e * ESTRIDE + 1];
          if (ty == KG_C && sb1 != 0.0) {
            double a = 2.0 * prm[e * KG_NPARAM] / h * sb1;
            if (p >= 0) { atomicAdd(&xbar[p], a); atomicAdd(&xb0[p], -a); }
            if (m >= 0) { atomicAdd(&xbar[m], -a); atomicAdd(&xb0[m], a); }
            double d1 = (p >= 0 ? x1[p] : 0.0) - (m >= 0 ? x1[m] : 0.0);
            double d0 = (p >= 0 ? xo[p] : 0.0) - (m >= 0 ? xo[m] : 0.0);
            M.ed[e * ESTRIDE + 2] += sb1 * 2.0 / h * (d1 - d0);
          } else if (ty == KG_L && sb1 != 0.0) {
            if (p >= 0) atomicAdd(&xbar[p], sb1);
            if (m >= 0) atomicAdd(&xbar[m], -sb1);
          }
        }
      __syncthreads();
      // lambda = A^-T xbar (rows of lambda are equations)
      if constexpr (std::is_same<F, A>::value) {
        if (act)
          for (int j = t; j < P.n; j += TPS) M.x[j] = xbar[j];
        __syncthreads();
        lu_solve_t<F, TPS>(P, M.v, M.x, t, act);
        if (act)
          for (int j = t; j < P.n; j += TPS) lam[j] = M.x[j];
        __syncthreads();
      } else {
        if (act)
          for (int j = t; j < P.n; j += TPS) M.w[j] = cvt<F>(xbar[j]);
        __syncthreads();
        lu_solve_t<F, TPS>(P, M.v, M.w, t, act);
        if (act)
          for (int j = t; j < P.n; j += TPS) lam[j] = re_of(M.w[j]);
        __syncthreads();
        for (int it = 0; it < aa.refine; it++) {  // FP64 refinement of the transposed solve
          if (act)
            for (int j = t; j < P.n; j += TPS) rr[j] = xbar[j];
          __syncthreads();
          if (act)
            for (int i = t; i < P.n; i += TPS)
              for (int p = P.row_ptr[i]; p < P.row_ptr[i + 1]; p++) {
                int q0 = P.scon_ptr[p], q1 = P.scon_ptr[p + 1];
                if (q0 == q1) continue;
                double a = 0.0;
                for (int q = q0; q < q1; q++) a += re_of(mat_contrib<A>(P.scon[q], P, prm, eb, sa, M.ed));
                atomicAdd(&rr[P.col[p]], -a * lam[i]);
              }
          __syncthreads();
          double rmax = 0.0;
          if (act)
            for (int j = t; j < P.n; j += TPS) rmax = fmax(rmax, fabs(rr[j]));
          double sc = sys_max<TPS>(rmax, red, ls, t);
          if (act && sc > 0.0)
            for (int j = t; j < P.n; j += TPS) M.w[j] = cvt<F>(rr[j] / sc);
          __syncthreads();
          lu_solve_t<F, TPS>(P, M.v, M.w, t, act && sc > 0.0);  // barriers are reached by every system
          if (act && sc > 0.0)
            for (int j = t; j < P.n; j += TPS) lam[j] += sc * re_of(M.w[j]);
          __syncthreads();
        }
      }
      // element terms of step s -> s+1
      if (act)
        for (int e = t; e < P.ne; e += TPS) {
          int ty = P.etype[e];
          if (ty != KG_R && ty != KG_C && ty != KG_L) continue;
          int p = P.ep[e], m = P.em[e];
          double v = prm[e * KG_NPARAM];
          if (ty == KG_L) {
            int c = P.ec[e], rc = P.erc[e];
            double lb = lam[rc];
            atomicAdd(&xb0[c], -al * (v / h) * lb);
            M.ed[e * ESTRIDE + 1] = trap ? -lb : 0.0;
            M.ed[e * ESTRIDE + 2] += lb * (-al / h) * xo[c] + al / h * lb * x1[c];
            continue;
          }
          double el = (P.erp[e] >= 0 ? lam[P.erp[e]] : 0.0) - (P.erm[e] >= 0 ? lam[P.erm[e]] : 0.0);
          double d1 = (p >= 0 ? x1[p] : 0.0) - (m >= 0 ? x1[m] : 0.0);
          if (ty == KG_R) {
            M.ed[e * ESTRIDE + 2] += el * d1 / (v * v);
          } else {
            double d0 = (p >= 0 ? xo[p] : 0.0) - (m >= 0 ? xo[m] : 0.0);
            double a = al * (v / h) * el;
            if (p >= 0) atomicAdd(&xb0[p], a);
            if (m >= 0) atomicAdd(&xb0[m], -a);
            M.ed[e * ESTRIDE + 1] = trap ? (-M.ed[e * ESTRIDE + 1] + el) : 0.0;
            M.ed[e * ESTRIDE + 2] += el * al / h * d0 - al / h * el * d1;
          }
        }
      __syncthreads();
      const double w = (s == 0) ? 0.5 * h : h;
      if (act)
        for (int j = t; j < P.n; j += TPS) {
          xbar[j] = xb0[j] + w * cvec[P.col_of[j]];
          xb0[j] = 0.0;
        }
      __syncthreads();
    }
  }
  // vl_0 = v(x_0): its adjoint flows into x_0
  if (act && trap)
    for (int e = t; e < P.ne; e += TPS) {
      if (P.etype[e] != KG_L) continue;
      double sb1 = M.ed[e * ESTRIDE + 1];
      int p = P.ep[e], m = P.em[e];
      if (p >= 0) atomicAdd(&xbar[p], sb1);
      if (m >= 0) atomicAdd(&xbar[m], -sb1);
    }
  __syncthreads();
  if (act) {
    double* xo = xbar0 + xop_off[ii];
    for (int j = t; j < P.n; j += TPS) xo[P.col_of[j]] = xbar[j];
    double* g = grad + (int64_t)ii * grad_stride;
    for (int e = t; e < P.ne; e += TPS) g[e] = M.ed[e * ESTRIDE + 2];
  }
  if (valid && t == 0) status[inst] = M.flags[0];
}

// Operating-point part of the adjoint: x_0 solves the DC system A_dc x_0 =
// b_dc, so dJ/dR gains (e^T lam0)(e^T x_0)/R^2 with lam0 = A_dc^-T xbar0.
// FP64; DC plan; xbar0 and xop in krets numbering.
template <int TPS>
__global__ void KG_LB(TPS) k_adj_op(BatchDev B, int kind_slot, const double* xbar0, const double* xop, const int64_t* xop_off,
                         double* grad, int64_t grad_stride, int32_t* status, Lay lay) {
  extern __shared__ __align__(16) char smem[];
  const int spb = blockDim.x / TPS;
  const int ls = threadIdx.x / TPS, t = threadIdx.x % TPS;
  const int inst = (int)(xcd_block() * spb + ls);
  const bool valid = inst < B.n_instances;
  const int ii = valid ? inst : 0;
  const PatDev& P = B.pats[B.inst_pat[kind_slot * B.n_instances + ii]];
  size_t sb = sys_bytes(lay);
  SysMem<double, double> M = carve<double, double>(smem + ls * sb, lay);
  const double* prm = B.params + B.eoff[ii] * KG_NPARAM;
  const double* eb = B.ebase + B.eoff[ii];
  if (t == 0) M.flags[0] = valid ? status[inst] : KG_OK;
  __syncthreads();
  bool act = valid && M.flags[0] == KG_OK;
  StampArgs sa{AK_DC, 0.0, 1.0, 1.0, 0.0, -1, 0.0};
  stamp_matrix<double, double, TPS>(P, M.v, prm, eb, sa, M.ed, t, act);
  if (act)
    for (int j = t; j < P.n; j += TPS) M.x[j] = xbar0[xop_off[ii] + P.col_of[j]];
  __syncthreads();
  factor<double, TPS>(P, M.v, t, act, &M.flags[0]);
  lu_solve_t<double, TPS>(P, M.v, M.x, t, act);
  if (act) {
    const double* x0 = xop + xop_off[ii];
    double* g = grad + (int64_t)ii * grad_stride;
    for (int e = t; e < P.ne; e += TPS) {
      if (P.etype[e] != KG_R) continue;
      int p = P.ep[e], m = P.em[e];
      double el = (P.erp[e] >= 0 ? M.x[P.erp[e]] : 0.0) - (P.erm[e] >= 0 ? M.x[P.erm[e]] : 0.0);
      double ex = (p >= 0 ? x0[P.col_of[p]] : 0.0) - (m >= 0 ? x0[P.col_of[m]] : 0.0);
      double v = prm[e * KG_NPARAM];
      g[e] += el * ex / (v * v);
    }
  }
  __syncthreads();
  if (valid && t == 0) status[inst] = M.flags[0];
}

// ------------------------------------------------------- launch helpers
template <class K, class... Args>
hipError_t launch(K fn, dim3 g, dim3 b, size_t lds, hipStream_t st, Args... args) {
  hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, g, b, lds, st, args...);
  return hipGetLastError();
}

#define KG_TPS_SWITCH(tps, CALL) \
  switch (tps) {                 \
    case 1: { constexpr int T = 1; return CALL; }     \
    case 4: { constexpr int T = 4; return CALL; }     \
    case 16: { constexpr int T = 16; return CALL; }   \
    case 64: { constexpr int T = 64; return CALL; }   \
    case 256: { constexpr int T = 256; return CALL; } \
    default: return hipErrorInvalidValue;             \
  }

hipError_t run_newton(int tps, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot, NewtonArgs na,
                      int sweep_e, double start, double step, int npts, double* out, const int64_t* out_off,
                      int32_t* status, int32_t* iters, Lay lay) {
  if (tps == 1024)  // one large system per CU: four times the threads per level of 256
    return launch(k_newton_dc<1024>, g, b, lds, st, B, kind_slot, na, sweep_e, start, step, npts, out, out_off,
                  status, iters, lay);
  if (tps == 512)  // the same with twice the registers per thread (256 VGPRs)
    return launch(k_newton_dc<512>, g, b, lds, st, B, kind_slot, na, sweep_e, start, step, npts, out, out_off,
                  status, iters, lay);
  KG_TPS_SWITCH(tps, launch(k_newton_dc<T>, g, b, lds, st, B, kind_slot, na, sweep_e, start, step, npts, out,
                            out_off, status, iters, lay))
}

hipError_t run_ac(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                  const double* freqs, int nf, const double* xop, const int64_t* xop_off, double* out,
                  const int64_t* out_off, int32_t* status, double* err, int refine, Lay lay,
                  unsigned long long* fixq, int64_t fixcap) {
  if (mixed) {
    KG_TPS_SWITCH(tps, launch(k_ac<c64, c128, T>, g, b, lds, st, B, kind_slot, freqs, nf, xop, xop_off, out,
                              out_off, status, err, refine, lay, fixq, fixcap))
  }
  KG_TPS_SWITCH(tps, launch(k_ac<c128, c128, T>, g, b, lds, st, B, kind_slot, freqs, nf, xop, xop_off, out, out_off,
                            status, err, refine, lay, fixq, fixcap))
}

hipError_t run_ac_fix(int grid, hipStream_t st, BatchDev B, int kind_slot, const double* freqs, int nf,
                      const double* xop, const int64_t* xop_off, double* out, const int64_t* out_off,
                      int32_t* status, double* err, int err_mode, const unsigned long long* fixq, int64_t fixcap,
                      int nmax, int ne_max, int nskip, int defer, char* gws, int64_t gstride) {
  return launch(k_ac_fix, dim3(grid), dim3(KG_FIX_NT), gws ? 0 : ac_fix_lds(nmax, ne_max), st, B, kind_slot, freqs,
                nf, xop, xop_off, out, out_off, status, err, err_mode, fixq, fixcap, nmax, ne_max, nskip, defer, gws,
                gstride);
}

hipError_t run_lte(int tps, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot, NewtonArgs na,
                   LteArgs la, const double* xop, const int64_t* xop_off, double* tout, double* xout,
                   const int64_t* xout_off, int32_t* npts, int32_t* newton_total, int32_t* rejected, int32_t* status,
                   Lay lay, int plan_ints) {
  if (plan_ints > 0) {
    KG_TPS_SWITCH(tps, launch(k_tran_lte<T, true>, g, b, lds, st, B, kind_slot, na, la, xop, xop_off, tout, xout,
                              xout_off, npts, newton_total, rejected, status, lay, plan_ints))
  }
  KG_TPS_SWITCH(tps, launch(k_tran_lte<T, false>, g, b, lds, st, B, kind_slot, na, la, xop, xop_off, tout, xout,
                            xout_off, npts, newton_total, rejected, status, lay, 0))
}

// Fill kernels: used instead of memset nodes so that a launch sequence can be
// captured into a hipGraph and replayed.
__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// Copies two word ranges (a -> da, b -> db) in one launch; used to publish a
// call's statuses and per-instance estimates into pinned host memory.
// Copies up to KG_COPY_SEGS word arrays in one launch (the results of a call
// into the pinned staging area).
__global__ void k_copy_segments(CopySegs c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < c.n; k++)
    if (i >= c.start[k] && i < c.start[k + 1]) c.dst[k][i - c.start[k]] = c.src[k][i - c.start[k]];
}
hipError_t run_copy_segments(const CopySegs& c, hipStream_t st) {
  const int64_t n = c.start[c.n];
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_segments, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c);
  return hipGetLastError();
}

hipError_t run_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
// Zeroes two arrays in one launch (one node fewer in a captured call graph).
__global__ void k_zero2_f64(double* a, int64_t na, double* b, int64_t nb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < na) a[i] = 0.0;
  else if (i < na + nb) b[i - na] = 0.0;
}
hipError_t run_zero2_f64(double* a, int64_t na, double* b, int64_t nb, hipStream_t st) {
  if (na + nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero2_f64, dim3((unsigned)((na + nb + 255) / 256)), dim3(256), 0, st, a, na, b, nb);
  return hipGetLastError();
}
hipError_t run_fill_f64(double* p, double v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

// plan_ints > 0 (64 or 256 threads per system): the plan is staged in LDS.
#define KG_STAGED_SWITCH(tps, CALL)                         \
  switch (tps) {                                            \
    case 64: { constexpr int T = 64; return CALL; }         \
    case 256: { constexpr int T = 256; return CALL; }       \
    default: return hipErrorInvalidValue;                   \
  }

hipError_t run_adj_fwd(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                       AdjArgs aa, const double* cvec, const double* xop, const int64_t* xop_off, double* ckpt,
                       int64_t ck_stride, int cpw, double* J, int32_t* status, Lay lay, int plan_ints) {
  if (plan_ints > 0) {
    if (mixed) {
      KG_STAGED_SWITCH(tps, launch(k_adj_fwd<float, double, T, true>, g, b, lds, st, B, kind_slot, aa, cvec, xop,
                                   xop_off, ckpt, ck_stride, cpw, J, status, lay, plan_ints))
    }
    KG_STAGED_SWITCH(tps, launch(k_adj_fwd<double, double, T, true>, g, b, lds, st, B, kind_slot, aa, cvec, xop,
                                 xop_off, ckpt, ck_stride, cpw, J, status, lay, plan_ints))
  }
  if (mixed) {
    KG_TPS_SWITCH(tps, launch(k_adj_fwd<float, double, T, false>, g, b, lds, st, B, kind_slot, aa, cvec, xop, xop_off,
                              ckpt, ck_stride, cpw, J, status, lay, 0))
  }
  KG_TPS_SWITCH(tps, launch(k_adj_fwd<double, double, T, false>, g, b, lds, st, B, kind_slot, aa, cvec, xop, xop_off,
                            ckpt, ck_stride, cpw, J, status, lay, 0))
}

hipError_t run_adj_bwd(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                       AdjArgs aa, const double* cvec, const double* ckpt, int64_t ck_stride, int cpw, double* seg,
                       int64_t seg_stride, double* grad, int64_t grad_stride, double* xbar0,
                       const int64_t* xop_off, int32_t* status, Lay lay, int plan_ints) {
  if (plan_ints > 0) {
    if (mixed) {
      KG_STAGED_SWITCH(tps, launch(k_adj_bwd<float, double, T, true>, g, b, lds, st, B, kind_slot, aa, cvec, ckpt,
                                   ck_stride, cpw, seg, seg_stride, grad, grad_stride, xbar0, xop_off, status, lay,
                                   plan_ints))
    }
    KG_STAGED_SWITCH(tps, launch(k_adj_bwd<double, double, T, true>, g, b, lds, st, B, kind_slot, aa, cvec, ckpt,
                                 ck_stride, cpw, seg, seg_stride, grad, grad_stride, xbar0, xop_off, status, lay,
                                 plan_ints))
  }
  if (mixed) {
    KG_TPS_SWITCH(tps, launch(k_adj_bwd<float, double, T, false>, g, b, lds, st, B, kind_slot, aa, cvec, ckpt,
                              ck_stride, cpw, seg, seg_stride, grad, grad_stride, xbar0, xop_off, status, lay, 0))
  }
  KG_TPS_SWITCH(tps, launch(k_adj_bwd<double, double, T, false>, g, b, lds, st, B, kind_slot, aa, cvec, ckpt,
                            ck_stride, cpw, seg, seg_stride, grad, grad_stride, xbar0, xop_off, status, lay, 0))
}

hipError_t run_adj_op(int tps, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                      const double* xbar0, const double* xop, const int64_t* xop_off, double* grad,
                      int64_t grad_stride, int32_t* status, Lay lay) {
  KG_TPS_SWITCH(tps, launch(k_adj_op<T>, g, b, lds, st, B, kind_slot, xbar0, xop, xop_off, grad, grad_stride, status,
                            lay))
}

hipError_t run_ac_persistent(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B,
                             int kind_slot, const double* freqs, int nf, const double* xop, const int64_t* xop_off,
                             double* out, const int64_t* out_off, int32_t* status, double* err, int refine, Lay lay,
                             int plan_ints, int64_t chunk, unsigned long long* fixq, int64_t fixcap) {
  if (mixed) {
    KG_TPS_SWITCH(tps, launch(k_ac_p<c64, c128, T>, g, b, lds, st, B, kind_slot, freqs, nf, xop, xop_off, out,
                              out_off, status, err, refine, lay, plan_ints, chunk, fixq, fixcap))
  }
  KG_TPS_SWITCH(tps, launch(k_ac_p<c128, c128, T>, g, b, lds, st, B, kind_slot, freqs, nf, xop, xop_off, out,
                            out_off, status, err, refine, lay, plan_ints, chunk, fixq, fixcap))
}

hipError_t run_tran(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                    NewtonArgs na, TranArgs ta, const double* xop, const int64_t* xop_off, double* out,
                    const int64_t* out_off, int32_t* status, int32_t* iters, double* err, Lay lay, int plan_ints) {
  if (plan_ints > 0) {  // staged plan: level-scheduled mappings only
    if (tps != 64 && tps != 256) return hipErrorInvalidValue;
    if (mixed) {
      if (tps == 64)
        return launch(k_tran<float, double, 64, true>, g, b, lds, st, B, kind_slot, na, ta, xop, xop_off, out, out_off,
                      status, iters, err, lay, plan_ints);
      return launch(k_tran<float, double, 256, true>, g, b, lds, st, B, kind_slot, na, ta, xop, xop_off, out, out_off,
                    status, iters, err, lay, plan_ints);
    }
    if (tps == 64)
      return launch(k_tran<double, double, 64, true>, g, b, lds, st, B, kind_slot, na, ta, xop, xop_off, out, out_off,
                    status, iters, err, lay, plan_ints);
    return launch(k_tran<double, double, 256, true>, g, b, lds, st, B, kind_slot, na, ta, xop, xop_off, out, out_off,
                  status, iters, err, lay, plan_ints);
  }
  if (mixed) {
    KG_TPS_SWITCH(tps, launch(k_tran<float, double, T, false>, g, b, lds, st, B, kind_slot, na, ta, xop, xop_off, out,
                              out_off, status, iters, err, lay, 0))
  }
  KG_TPS_SWITCH(tps, launch(k_tran<double, double, T, false>, g, b, lds, st, B, kind_slot, na, ta, xop, xop_off, out,
                            out_off, status, iters, err, lay, 0))
}

hipError_t run_tran_lin(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                        const double* xop, const int64_t* xop_off, double* out, const int64_t* out_off,
                        int32_t* status, int32_t* iters, double* err, double* srcv, int nsrc_max) {
  const dim3 b(tps);
#define KG_LIN(T, RR)                                                                                          \
  if (tps == T && r == RR)                                                                                   \
    return launch(k_tran_lin<T, RR>, g, b, lds, st, B, kind_slot, ta, xop, xop_off, out, out_off, status, iters, \
                  err, srcv, nsrc_max);
  KG_LIN(64, 1) KG_LIN(64, 2) KG_LIN(64, 4)
  KG_LIN(128, 1) KG_LIN(128, 2) KG_LIN(128, 4)
  KG_LIN(256, 1) KG_LIN(256, 2) KG_LIN(256, 4)
  KG_LIN(512, 1) KG_LIN(512, 2) KG_LIN(512, 4)
  KG_LIN(1024, 1) KG_LIN(1024, 2)
#undef KG_LIN
  return hipErrorInvalidValue;
}

hipError_t run_adj_lin_fwd(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                           int K, const double* cvec, const double* xop, const int64_t* xop_off, double* ckpt,
                           int64_t ck_stride, int cpw, double* J, int32_t* status, double* srcv, int nsrc_max) {
  const dim3 b(tps);
#define KG_LIN(T, RR)                                                                                            \
  if (tps == T && r == RR)                                                                                     \
    return launch(k_adj_lin_fwd<T, RR>, g, b, lds, st, B, kind_slot, ta, K, cvec, xop, xop_off, ckpt, ck_stride, \
                  cpw, J, status, srcv, nsrc_max);
  KG_LIN(64, 1) KG_LIN(64, 2) KG_LIN(64, 4)
  KG_LIN(256, 1) KG_LIN(256, 2) KG_LIN(256, 4)
  KG_LIN(512, 1) KG_LIN(512, 2) KG_LIN(512, 4)
#undef KG_LIN
  return hipErrorInvalidValue;
}

hipError_t run_adj_lin_rec(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                           int K, const double* ckpt, int64_t ck_stride, int cpw, double* traj, int64_t traj_stride,
                           const int32_t* status, double* srcv, int nsrc_max) {
  const dim3 b(tps);
#define KG_LIN(T, RR)                                                                                             \
  if (tps == T && r == RR)                                                                                      \
    return launch(k_adj_lin_rec<T, RR>, g, b, lds, st, B, kind_slot, ta, K, ckpt, ck_stride, cpw, traj, traj_stride, \
                  status, srcv, nsrc_max);
  KG_LIN(64, 1) KG_LIN(64, 2) KG_LIN(64, 4)
  KG_LIN(256, 1) KG_LIN(256, 2) KG_LIN(256, 4)
  KG_LIN(512, 1) KG_LIN(512, 2) KG_LIN(512, 4)
#undef KG_LIN
  return hipErrorInvalidValue;
}

hipError_t run_adj_lin_bwd(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                           int K, const double* cvec, const double* ckpt, int64_t ck_stride, int cpw, double* seg,
                           int64_t seg_stride, double* grad, int64_t grad_stride, double* xbar0,
                           const int64_t* xop_off, int32_t* status, double* srcv, int nsrc_max,
                           const double* xs_all, int64_t xs_inst_stride, int64_t xs_step_stride) {
  const dim3 b(tps);
#define KG_LIN(T, RR)                                                                                        \
  if (tps == T && r == RR)                                                                                 \
    return launch(k_adj_lin_bwd<T, RR>, g, b, lds, st, B, kind_slot, ta, K, cvec, ckpt, ck_stride, cpw, seg, \
                  seg_stride, grad, grad_stride, xbar0, xop_off, status, srcv, nsrc_max, xs_all, xs_inst_stride,  \
                  xs_step_stride);
  KG_LIN(64, 1) KG_LIN(64, 2) KG_LIN(64, 4)
  KG_LIN(256, 1) KG_LIN(256, 2) KG_LIN(256, 4)
  KG_LIN(512, 1) KG_LIN(512, 2) KG_LIN(512, 4)
#undef KG_LIN
  return hipErrorInvalidValue;
}

}  // namespace kg
