// Pattern-specialised kernels compiled at run time (hiprtc, gfx950).
//
// For a shared topology whose LU fits in one lane's registers, the host
// symbolic plan is turned into straight-line HIP source: every LU slot, every
// solution entry and every element-value load has a fixed register or a fixed
// offset, so one lane solves one system with no plan metadata traffic, no
// LDS and no barriers. The generated source is compiled once per
// (device, source text) and cached for the life of the process.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "kg_internal.h"

namespace kg {

struct JitFn {
  hipFunction_t fn = nullptr;
  int vgprs = 0;         // registers per lane reported by the loader
  int scratch = 0;       // private (spill) bytes per lane
};

// Safety factor of the mixed precision kernel's adaptive stop: the error left
// after a refinement pass is predicted as KG_REFINE_SAFETY * (|dx| / |x|)^2
// (max norms over the components, max_i max(|re_i|, |im_i|)). On the batched
// RC ladders (2 x 16 instances x 200 frequencies) the observed ratio of the
// true normwise error to that squared ratio stays below 130 (median 3).
#define KG_REFINE_SAFETY 200.0
// Default target of that prediction (the kernel's `stop` argument; 0 runs
// every pass): two orders below the 1e-9 accuracy the mixed path promises.
constexpr double KG_REFINE_STOP = 1e-11;

// Source of the lane-per-system AC sweep kernel (jit_ac_name) for one pattern
// (AC kind, shared by every instance); the refinement count of the mixed
// precision kernel and its adaptive stop are launch arguments (refine: the
// most passes; stop > 0: a wave leaves once every lane's predicted error is
// at most stop). xsplit (mixed precision only)
// keeps the solution as x0 + dx: the complex64 first solution x0 in LDS
// (512 * n bytes per 64-lane block, launch argument) and the complex64 sum
// of the corrections dx in registers, instead of an FP64 x in registers. Returns false, with the reason in `why`, when the pattern is
// outside what the generator handles (diodes, patterns whose per-lane state
// would not fit in registers).
bool jit_ac_source(const PatLayout& L, const int32_t* buf, bool mixed, bool xsplit, std::string& src,
                   std::string& why);

// Kernel symbol of the generated AC kernel: kg_ac_lane_c128 (FP64),
// kg_ac_lane_c64 (mixed, FP64 solution in registers), kg_ac_lane_c64_split
// (mixed, split solution).
const char* jit_ac_name(bool mixed, bool xsplit);

// Source of the lane-per-instance LTE-controlled backward-Euler transient
// kernel (jit_lte_name) for one AK_TRAN pattern shared by every instance:
// the step controller, Newton loop, junction limiting and stamps of the
// generic k_tran_lte with the plan written out as straight-line code.
// Returns false, with the reason in `why`, when the per-lane state would not
// fit in registers.
bool jit_lte_source(const PatLayout& L, const int32_t* buf, std::string& src, std::string& why);
const char* jit_lte_name();

// Source of the lane-per-instance fixed-step transient kernel (jit_tran_name)
// for one AK_TRAN pattern shared by every instance: BE-compat, BE and TRAP
// (launch argument), the Newton loop, companions and stamps of k_tran. FP64.
bool jit_tran_source(const PatLayout& L, const int32_t* buf, std::string& src, std::string& why);
const char* jit_tran_name();

// Source of the lane-per-instance kernel for linear operating points and DC
// sweeps (jit_dc_name) for one AK_DC pattern shared by every instance: one
// straight-line solve per point. Declines nonlinear circuits.
bool jit_dc_source(const PatLayout& L, const int32_t* buf, std::string& src, std::string& why);
const char* jit_dc_name();

// Compiles `src` for the device's architecture and loads kernel `name`;
// cached per (device, source).
hipError_t jit_load(int device, const std::string& src, const char* name, JitFn& out, std::string& log);

}  // namespace kg
