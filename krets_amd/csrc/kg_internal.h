// Internal definitions shared by the host symbolic pass and the gfx950 kernels.
#pragma once
#include <cstdint>

#include "krets_gpu.h"

namespace kg {

// Which of krets' stamp families a pattern is built for. DC covers op.rs and
// dc.rs (capacitors filtered out, op.rs:20-24), AC covers ac.rs, TRAN covers
// transient.rs (backward-Euler companions) and the trapezoidal additions.
enum AnalysisKind : int32_t { AK_DC = 0, AK_AC = 1, AK_TRAN = 2, AK_COUNT = 3 };

// Symmetric fill-reducing orders applied after the structural matching.
// ORD_BALANCED: minimum fill with ties to the shallowest elimination-tree
// position (independent elimination chains for one thread per system).
enum Ordering : int32_t { ORD_NATURAL = 0, ORD_MINDEG = 1, ORD_NESTED = 2, ORD_BALANCED = 3 };

// Matrix contribution kinds gathered per LU slot. A contribution is encoded as
// (element << 3) | kind.
enum : int32_t {
  CK_YP = 0,    // +y(e): 1/R, j*w*C, alpha*C/h, diode G
  CK_YN = 1,    // -y(e)
  CK_ONEP = 2,  // +1 (incidence, current-source diagonal, G2 capacitor AC diagonal)
  CK_ONEN = 3,  // -1
  CK_BR = 4,    // branch diagonal: -j*w*L, -alpha*L/h, -R (G2 resistor)
  CK_KP = 5,    // +k(e): controlled-source gain (E, G, F, H)
  CK_KN = 6     // -k(e)
};

// Right-hand-side contribution kinds gathered per permuted row.
enum : int32_t {
  RK_SRC = 0,  // independent source value (V dc/ac/waveform, I dc)
  RK_DP = 1,   // -Ieq of a diode at its plus node
  RK_DM = 2,   // +Ieq at its minus node
  RK_CP = 3,   // +capacitor companion current at plus node
  RK_CM = 4,   // -capacitor companion current at minus node
  RK_IND = 5   // inductor companion on its branch row
};

// Accuracy guard of the AC sweeps. Pivots are fixed once per topology by the
// symbolic pass, at one representative frequency and from one instance, so a
// solve far from that point can lose backward stability (an inductor branch
// pivot of magnitude w*L chosen at the sweep's centre is tiny at its low end).
//
// Lane kernels (jit.cpp) bound it a priori: during the factorisation they
// form rho = || |L||U| ||_inf / ||A||_inf (entry norm |re| + |im|, |L|
// bounded through |l_ik| <= |v_ik| |1/u_kk|). LU without pivoting and the two
// triangular solves give (A + dA) x = b with |dA| <= (3 g_m + g_m^2) |L||U|
// (Higham, Accuracy and Stability of Numerical Algorithms, Thm 9.4, with the
// complex g_m = m sqrt(2) gamma_2 / (1 - m sqrt(2) gamma_2), m the longest
// row of the factor pattern), so the normwise backward error is at most
// guard_bound_coeff(m) * rho. rho above KG_GUARD_GROWTH queues the solve; the
// instance's FP64 figure is the largest bound. A mixed precision solve is
// queued when the growth of its complex64 factorisation is above
// KG_GUARD_GROWTH (refinement through factors that grew can settle short of
// FP64 accuracy with small corrections) or its refinement estimate stays
// above KG_GUARD_MIXED.
// Level-scheduled kernels (k_ac, k_ac_p) measure instead, in FP64 from the
// LDS-resident b and x, the componentwise backward error (Oettli-Prager)
//   omega = max_i |b - A x|_i / ((|A| |x| + |b|)_i + u ||b||_inf),
// which bounds the normwise one, against KG_GUARD_OMEGA (the floor
// u ||b||_inf keeps rows whose solution decays into the subnormal range
// measurable), and for mixed precision the returned solution's normwise
// backward error against KG_GUARD_MIXED_ETA.
// Queued solves (and failed static pivots) are solved again by k_ac_fix: a
// dense LU with row partial pivoting in FP64 (systems of at most KG_FIX_NMAX
// unknowns), then up to two FP64 refinement steps, after which its normwise
// backward error eta is measured; eta above KG_FIX_FAIL reports
// KG_DECOMPOSITION, and eta is the figure of a fixed solve. Larger queued
// systems (up to KG_FIX_GMAX unknowns) take the same dense LU in a global
// workspace region per block, in a second launch.
constexpr double KG_GUARD_GROWTH = 4.0;
constexpr double KG_GUARD_OMEGA = 1.4210854715202004e-14;  // 2^-46: 64 units in the last place
constexpr double KG_GUARD_MIXED = 1e-10;                  // a tenth of the 1e-9 the mixed path promises
constexpr double KG_GUARD_MIXED_ETA = 1.4210854715202004e-14;  // 2^-46, normwise, on the returned solution
// guard_bound_coeff(m) = 3.01 sqrt(2) g_m (the sqrt(2): rho is measured in
// |re| + |im| norms, the bound holds in moduli)
#ifdef __HIP__
#define KG_HOST_DEVICE __host__ __device__
#else  // the host-only sanitizer build of the symbolic pass (g++)
#define KG_HOST_DEVICE
#endif
KG_HOST_DEVICE inline double guard_bound_coeff(int m) {
  const double u = 0x1p-53, g2 = 2.0 * u / (1.0 - 2.0 * u), q = m * 1.4142135623730951 * g2;
  return 3.01 * 1.4142135623730951 * q / (1.0 - q);
}
// Static-pivot guard of the FP64 OP, DC sweep and transient factorisations
// (level-scheduled and lane kernels): the same a-priori bound,
// guard_bound_coeff(m) * rho, formed from each returned solution's factors
// (every Newton iterate of a nonlinear circuit that ends a point or a step;
// once per instance for a linear transient's single factorisation). There
// is no fixup: a bound above KG_GUARD_STATIC reports KG_DECOMPOSITION (krets'
// partial pivoting would have re-pivoted), and a transient's FP64 err_est is
// the largest bound of the instance.
constexpr double KG_GUARD_STATIC = 9.313225746154785e-10;  // 2^-30
constexpr double KG_FIX_FAIL = 9.094947017729282e-13;     // 2^-40
constexpr int KG_FIX_NMAX = 90;                           // 90 x 90 complex128 in LDS: 127 KB
constexpr int KG_FIX_GMAX = 1024;  // larger queued systems: a dense LU in a global region per block (16 MB)
constexpr int KG_FIX_GBLOCKS = 16;  // blocks (regions) of that second launch
// Fixup queue layout (unsigned 64-bit words): [0] count, [1 + q] queued system
// (instance * number of frequencies + frequency index), q < capacity.

// Layout of one pattern inside a flat int32 buffer. Offsets are in int32 units.
struct PatLayout {
  int32_t n, nnz, ne, nlev, nblev, ntlev, ntblev, nonlinear;
  int32_t has_diode, has_dyn;  // diode present; C or L present (companions)
  int32_t o_row_ptr, o_col, o_diag;
  int32_t o_fprog_ptr, o_fprog;
  int32_t o_lvl_ptr, o_lvl_rows, o_blvl_ptr, o_blvl_rows;
  int32_t o_scon_ptr, o_scon, o_rcon_ptr, o_rcon;
  int32_t o_col_of, o_cpos, o_isbr;
  int32_t o_etype, o_ep, o_em, o_ec;
  int32_t o_ut_ptr, o_ut, o_lt_ptr, o_lt;
  int32_t o_tlvl_ptr, o_tlvl_rows, o_tblvl_ptr, o_tblvl_rows;
  int32_t o_row_of, o_rpos, o_erp, o_erm, o_erc;
  int32_t status;  // KG_OK or KG_DECOMPOSITION (structurally singular)
  int32_t total;   // buffer length
};

// Compact plan of a linear transient (no diodes), built once per topology so
// that a system's whole step loop runs from LDS: 16-bit indices, one packed
// row descriptor per row and triangular solve (in level order), right-hand
// side gather lists and the element records the companion models need. The
// header is followed by uint16 arrays at the given word offsets; the buffer
// is padded to 16 bytes.
//   fdesc / bdesc (4 words per row, level order): row, first entry, last
//     entry + 1, diagonal entry (forward: strictly lower entries; backward:
//     strictly upper entries)
//   dyn entry (LIN_DYN words): (kind << 14) | element (kind 1: C, 2: L),
//     plus column, minus column, branch column (LIN_NONE for ground / none)
//   src entry (LIN_SRC words): kind (KG_V / KG_I), element
//   rcon code: (index << 2) | LK_* (index into src or dyn)
// After the first words_fwd words (all a transient needs) follows what the
// discrete adjoint adds:
//   tfdesc / tbdesc (4 words per row, transposed-solve level order): row,
//     first entry, end entry, diagonal (U^T: strictly lower entries of
//     column i of U; L^T: of L), entries in tent as (column, LU slot) pairs
//   adj entry (LIN_ADJ words): (kind << 14) | element (kind 0: R, 1: C,
//     2: L), plus / minus column, plus / minus row (KCL equations), branch
//     column, branch row, companion (dyn) index (LIN_NONE where absent)
constexpr int LIN_DYN = 4, LIN_SRC = 2, LIN_ADJ = 8;
constexpr uint16_t LIN_NONE = 0xFFFF;
enum LinCode { LK_SRC = 0, LK_CP = 1, LK_CM = 2, LK_IND = 3 };
struct LinHdr {
  int32_t n, nnz, nlev, nblev, ndyn, ncon, nsrc, words;
  int32_t o_lvl_ptr, o_fdesc, o_blvl_ptr, o_bdesc, o_col, o_cpos;
  int32_t o_rcon_ptr, o_rcon, o_dyn, o_src;
  int32_t lev0_free;  // the first forward level has no lower-triangle entries
  int32_t words_fwd;  // prefix a forward transient copies
  int32_t ntlev, ntblev, o_tlvl_ptr, o_tfdesc, o_tblvl_ptr, o_tbdesc, o_tent, nadj, o_adj;
  int32_t tlev0_free;  // the first level of each transposed solve has no entries
  int32_t ne;          // elements (gradient slots)
  int32_t pad;
};

// Device view of one pattern (pointers into uploaded buffer).
struct PatDev {
  int32_t n, nnz, ne, nlev, nblev, ntlev, ntblev, nonlinear, has_diode, has_dyn;
  const int32_t *row_ptr, *col, *diag;
  const int32_t *fprog_ptr, *fprog;
  const int32_t *lvl_ptr, *lvl_rows, *blvl_ptr, *blvl_rows;
  const int32_t *scon_ptr, *scon, *rcon_ptr, *rcon;
  const int32_t *col_of, *cpos, *isbr;
  const int32_t *etype, *ep, *em, *ec;
  const int32_t *ut_ptr, *ut, *lt_ptr, *lt;
  const int32_t *tlvl_ptr, *tlvl_rows, *tblvl_ptr, *tblvl_rows;
  // equation (row) positions: rpos[original row] and, per element, the
  // permuted rows of its plus / minus KCL equations and branch equation
  const int32_t *row_of, *rpos, *erp, *erm, *erc;
  // the plan buffer itself (all pointers above point into it)
  const int32_t* base;
  int32_t total;
  // compact 16-bit plan of a linear transient (LinHdr + arrays) or null
  const uint16_t* lin;
  int32_t lin_words;
};

}  // namespace kg
