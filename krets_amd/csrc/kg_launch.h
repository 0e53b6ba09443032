// Launch interface between the host API (api.cpp) and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "kg_internal.h"

namespace kg {

// Element state stride in LDS (doubles per element): [0] diode G or
// companion state (capacitor current / inductor voltage), [1] diode Ieq (AC:
// the diode's small-signal capacitance), [2] diode junction voltage of the
// previous iterate, [3] trapezoidal junction-charge current of the previous
// step (diodes with junction charge).
constexpr int ESTRIDE = 4;

// Per-instance views.
struct BatchDev {
  const PatDev* pats;      // pattern table
  const int32_t* inst_pat; // instance -> pattern index (per analysis kind table)
  const double* params;    // [elements][KG_NPARAM]
  const double* ebase;     // [elements]: 1/R for resistors, else param[0]
  const int64_t* eoff;     // instance -> first element record
  const int32_t* n_inst;   // instance -> unknown count
  int32_t n_instances;
  // optional subset of instances for this launch (size-class buckets of an
  // uneven batch): slot k of the grid runs instance inst_list[k]
  const int32_t* inst_list = nullptr;
  int32_t n_list = 0;
};

// LDS carve-up of one system. The same layout is computed on the host
// (sys_bytes) to size the dynamic LDS of a launch.
struct Lay {
  int32_t nnz_max, n_max, ne_max;
  int32_t need_g, need_xp, need_ed;  // need_xp: number of extra FP64 n-vectors
  int32_t fsz, asz, mixed;
  int32_t stage_prm;  // the system's element records (params, value table) copied into LDS
  // k_newton_dc: ints of the plan prefix (factor and solve structure, up to
  // o_col) copied into LDS behind the block's systems and reduction buffer;
  // 0: read from global memory
  int32_t stage_plan = 0;
  // Systems too large for LDS: with gws set, the launch runs one system per
  // block, block b's system (values and reduction buffer) lives at
  // gws + b * gstride in global memory, and the launch covers systems
  // sys_off .. sys_off + gridDim.x - 1 (the kernels that take it: k_newton_dc,
  // k_ac, k_tran, k_tran_lte).
  char* gws = nullptr;
  int64_t gstride = 0, sys_off = 0;
};

__host__ __device__ inline size_t lay_al(size_t b) { return (b + 15) & ~size_t(15); }

#ifdef __HIPCC__
// Where a block's systems live: its LDS, or its region of the global workspace.
// The block's barriers order its accesses there as in LDS (one CU, one L1).
__device__ inline char* sys_region(char* smem, const Lay& lay) {
  return lay.gws ? lay.gws + (int64_t)blockIdx.x * lay.gstride : smem;
}
#endif

__host__ __device__ inline size_t sys_bytes(const Lay& L) {
  size_t s = 64 + lay_al((size_t)L.fsz * L.nnz_max) + 2 * lay_al((size_t)L.asz * L.n_max);
  if (L.mixed) s += lay_al((size_t)L.fsz * L.n_max) + lay_al((size_t)L.asz * L.n_max);
  if (L.need_g) s += lay_al(8 * (size_t)L.n_max);
  s += (size_t)L.need_xp * lay_al(8 * (size_t)L.n_max);
  if (L.need_ed) s += lay_al(8 * (size_t)ESTRIDE * L.ne_max);
  if (L.stage_prm) s += lay_al(8 * (size_t)(KG_NPARAM + 1) * L.ne_max);
  return s;
}

// LDS bytes of one system of the linear transient kernel (k_tran_lin): the
// plan prefix, factors, two solution buffers, companion states and
// coefficients (each with a trailing zero slot), source records, flags.
__host__ __device__ inline size_t lin_bytes(const LinHdr& h) {
  return lay_al(2 * (size_t)h.words_fwd) + 8 * ((size_t)h.nnz + 2 * (size_t)h.n + 2 * (size_t)h.ndyn + 5) +
         8 * (size_t)KG_NPARAM * h.nsrc + 64;
}

// LDS bytes of one system of the linear adjoint kernels (k_adj_lin_fwd /
// k_adj_lin_bwd): the whole plan, the transient's arrays, four more solution
// vectors (adjoint load and its next-step part, two recomputed states), the
// companion adjoint states, per-element gradients and a reduction buffer.
__host__ __device__ inline size_t lin_adj_bytes(const LinHdr& h) {
  return lay_al(2 * (size_t)h.words) + 8 * ((size_t)h.nnz + 2 * (size_t)h.n + 2 * (size_t)h.ndyn + 5) +
         8 * (size_t)KG_NPARAM * h.nsrc + 8 * (4 * ((size_t)h.n + 1) + (size_t)h.ndyn + 1 + (size_t)h.ne + 1 + 32) +
         64;
}

struct NewtonArgs {
  double reltol, iabstol, vabstol;
  int maxiter, compat;
};

struct TranArgs {
  double h, alpha;
  int nsteps, method;  // kg_method
  int refine;
};

// Discrete adjoint of a fixed-step linear transient.
struct AdjArgs {
  double h, alpha;
  int nsteps, method, refine, K, trap;
};

// Backward Euler with local-truncation-error control (clean mode).
struct LteArgs {
  double tstop, h0, hmin, hmax, lte_reltol, lte_abstol;
  int max_points;
};


hipError_t run_newton(int tps, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot, NewtonArgs na,
                      int sweep_e, double start, double step, int npts, double* out, const int64_t* out_off,
                      int32_t* status, int32_t* iters, Lay lay);
// AC sweep kernels. fixq / fixcap: the accuracy guard's queue (see the
// guard constants in kg_internal.h); null: no guard (statuses and error
// figures as computed).
hipError_t run_ac(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                  const double* freqs, int nf, const double* xop, const int64_t* xop_off, double* out,
                  const int64_t* out_off, int32_t* status, double* err, int refine, Lay lay,
                  unsigned long long* fixq, int64_t fixcap);
// Re-solves the systems in the guard's queue (k_ac_fix), persistent over the
// queue with `grid` blocks; err_mode 0: backward error, 1: correction size.
// Systems of nskip < n <= nmax unknowns, in LDS (gws null) or in a global
// workspace region per block (gws + block * gstride, ac_fix_lds bytes each);
// defer: larger systems are left to a later launch instead of failing.
size_t ac_fix_lds(int nmax, int ne_max);
hipError_t run_ac_fix(int grid, hipStream_t st, BatchDev B, int kind_slot, const double* freqs, int nf,
                      const double* xop, const int64_t* xop_off, double* out, const int64_t* out_off,
                      int32_t* status, double* err, int err_mode, const unsigned long long* fixq, int64_t fixcap,
                      int nmax, int ne_max, int nskip, int defer, char* gws, int64_t gstride);
hipError_t run_lte(int tps, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot, NewtonArgs na,
                   LteArgs la, const double* xop, const int64_t* xop_off, double* tout, double* xout,
                   const int64_t* xout_off, int32_t* npts, int32_t* newton_total, int32_t* rejected, int32_t* status,
                   Lay lay, int plan_ints = 0);
hipError_t run_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t st);
// Word-array copies of one launch: segment k copies start[k + 1] - start[k]
// words from src[k] to dst[k].
constexpr int KG_COPY_SEGS = 8;
struct CopySegs {
  const uint32_t* src[KG_COPY_SEGS];
  uint32_t* dst[KG_COPY_SEGS];
  int64_t start[KG_COPY_SEGS + 1];
  int n;
};
hipError_t run_copy_segments(const CopySegs& c, hipStream_t st);
hipError_t run_fill_f64(double* p, double v, int64_t n, hipStream_t st);
hipError_t run_zero2_f64(double* a, int64_t na, double* b, int64_t nb, hipStream_t st);
hipError_t run_adj_fwd(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                       AdjArgs aa, const double* cvec, const double* xop, const int64_t* xop_off, double* ckpt,
                       int64_t ck_stride, int cpw, double* J, int32_t* status, Lay lay, int plan_ints = 0);
hipError_t run_adj_bwd(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                       AdjArgs aa, const double* cvec, const double* ckpt, int64_t ck_stride, int cpw, double* seg,
                       int64_t seg_stride, double* grad, int64_t grad_stride, double* xbar0,
                       const int64_t* xop_off, int32_t* status, Lay lay, int plan_ints = 0);
hipError_t run_adj_op(int tps, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                      const double* xbar0, const double* xop, const int64_t* xop_off, double* grad,
                      int64_t grad_stride, int32_t* status, Lay lay);
hipError_t run_ac_persistent(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B,
                             int kind_slot, const double* freqs, int nf, const double* xop, const int64_t* xop_off,
                             double* out, const int64_t* out_off, int32_t* status, double* err, int refine, Lay lay,
                             int plan_ints, int64_t chunk, unsigned long long* fixq, int64_t fixcap);
// Linear fixed-step transient from LDS (k_tran_lin), one system per block of
// tps threads (64, 128, 256, 512, 1024), each owning up to r rows (1, 2, 4;
// 1024 threads: 1, 2); every launched system needs n <= r * tps and
// lds >= its lin_bytes. srcv: workspace of n_instances * nsrc_max *
// (nsteps + 1) doubles for the tabulated source values.
hipError_t run_tran_lin(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                        const double* xop, const int64_t* xop_off, double* out, const int64_t* out_off,
                        int32_t* status, int32_t* iters, double* err, double* srcv, int nsrc_max);
// Discrete adjoint from LDS (k_adj_lin_fwd / k_adj_lin_bwd), shared topology,
// one system per block of tps threads (64, 256, 512) owning up to r rows
// (1, 2, 4); lds >= lin_adj_bytes; checkpoints hold the permuted solution
// then the companion states (cpw >= n + ndyn).
hipError_t run_adj_lin_fwd(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                           int K, const double* cvec, const double* xop, const int64_t* xop_off, double* ckpt,
                           int64_t ck_stride, int cpw, double* J, int32_t* status, double* srcv, int nsrc_max);
// Backward pass. xs_all (optional): every state, [instance] (stride
// xs_inst_stride) [step] (stride xs_step_stride), permuted order; the pass
// then recomputes nothing.
hipError_t run_adj_lin_bwd(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                           int K, const double* cvec, const double* ckpt, int64_t ck_stride, int cpw, double* seg,
                           int64_t seg_stride, double* grad, int64_t grad_stride, double* xbar0,
                           const int64_t* xop_off, int32_t* status, double* srcv, int nsrc_max,
                           const double* xs_all = nullptr, int64_t xs_inst_stride = 0, int64_t xs_step_stride = 0);
// Recomputes every checkpoint segment in parallel (grid: instances x
// segments) into traj ([instance][nsteps + 1][n], permuted order, stride
// traj_stride per instance); lds >= lin_bytes.
hipError_t run_adj_lin_rec(int tps, int r, dim3 g, size_t lds, hipStream_t st, BatchDev B, int kind_slot, TranArgs ta,
                           int K, const double* ckpt, int64_t ck_stride, int cpw, double* traj, int64_t traj_stride,
                           const int32_t* status, double* srcv, int nsrc_max);
hipError_t run_tran(int tps, int mixed, dim3 g, dim3 b, size_t lds, hipStream_t st, BatchDev B, int kind_slot,
                    NewtonArgs na, TranArgs ta, const double* xop, const int64_t* xop_off, double* out,
                    const int64_t* out_off, int32_t* status, int32_t* iters, double* err, Lay lay,
                    int plan_ints = 0);

}  // namespace kg
