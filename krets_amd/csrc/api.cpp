// Host side of the C ABI: contexts, batch upload, the per-topology symbolic
// plan cache, launch configuration and the analysis entry points. Mirrors
// krets_solver::Solver's dispatch (crates/krets-solver/src/solver.rs:17-53)
// over a whole batch of instances.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "kg_jit.h"
#include "kg_launch.h"
#include "krets_gpu.h"

namespace kg {
std::vector<int32_t> build_pattern(const kg_element* el, int ne, int n, AnalysisKind ak, Ordering ord,
                                   const double* rep, double s_rep, PatLayout& L, std::string& err);
bool build_lin_plan(const PatLayout& L, const std::vector<int32_t>& b, std::vector<uint16_t>& out);
}

using namespace kg;

namespace {

struct PatEntry {
  PatLayout lay;
  std::vector<int32_t> host;  // the plan buffer (host copy)
  int32_t* d_buf = nullptr;
  std::vector<uint16_t> lin;  // compact linear-transient plan (empty if none)
  uint16_t* d_lin = nullptr;
  PatDev dev{};
  std::string err;
};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

constexpr size_t LDS_MAX = 160 * 1024;

// Device scratch of the allocation-free AC sweep (kg_ac_sweep_noalloc): one
// static array per GPU, loaded with the library's code object and held until
// the process ends. Host code never passes `scratch` itself to a kernel (on
// the host the name is only a stand-in): kg_create resolves its address on
// each device once, with that device current, and calls use that address.
constexpr size_t KG_NOALLOC_SCRATCH_BYTES = size_t(8) << 20;
constexpr int64_t KG_NOALLOC_FIXQ = 65536;  // accuracy-guard queue entries reserved in it
__device__ __align__(256) char scratch[KG_NOALLOC_SCRATCH_BYTES];
constexpr int KG_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[KG_MAX_DEVICES] = {};
hipError_t g_scratch_err[KG_MAX_DEVICES] = {};
bool g_scratch_done[KG_MAX_DEVICES] = {};

// Resolves the scratch array's address on `device` (current), once.
void scratch_init(int device) {
  if (device < 0 || device >= KG_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();
  g_scratch_done[device] = true;
}
constexpr size_t LDS_SHARED_BUDGET = 64 * 1024;

}  // namespace

struct kg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> aux;  // extra streams (default priority) for concurrent buckets
  std::string err;
  // batch
  int n_inst = 0;
  bool shared = false;
  std::vector<kg_element> el;
  std::vector<int64_t> eoff;
  std::vector<int32_t> nunk;
  std::vector<double> params, ebase;
  std::vector<int> group_of, group_first;
  double* d_params = nullptr;
  double* d_ebase = nullptr;
  // shared topologies: the same values in blocks of 64 instances, element
  // by element (entry (i / 64 * ne + e) * 64 + i % 64), read by the mixed
  // precision lane AC kernels, whose waves hold 64 consecutive instances
  double* d_ebase_t = nullptr;
  int64_t* d_eoff = nullptr;
  int32_t* d_nunk = nullptr;
  // plans
  std::map<std::tuple<int, int, int, double>, int> pat_index;
  std::vector<PatEntry> pats;
  PatDev* d_pats = nullptr;
  size_t d_pats_n = 0;
  std::map<std::tuple<int, int, double>, int32_t*> inst_pat;
  // run-time specialised kernels per (pattern, mixed, x in memory)
  std::map<std::tuple<int, int, int>, JitFn> lane_fn;
  std::map<std::string, DevBuf> ws;
  // last host contents copied into a workspace buffer (skips re-uploads)
  std::map<std::string, std::pair<void*, std::vector<char>>> uploaded;
  // adjoint state kept between kg_adjoint_forward and kg_adjoint_backward
  struct {
    bool ready = false;
    AdjArgs aa{};
    int tps = 0, mixed = 0, ncp = 0, cpw = 0;
    int32_t refine = 2;
    Lay lay{};
  } adj;
  // autotuner: best threads-per-system per workload key
  std::map<std::string, int> tune;
  std::string tune_file;
  std::vector<std::pair<int, float>> last_trials;
  // captured launch sequences of repeated calls (AC sweeps): key -> graph
  bool graphs = true;
  std::map<std::string, std::pair<hipGraph_t, hipGraphExec_t>> gexec;
  // pinned host staging for the small per-call outputs (statuses, error
  // estimates): a DMA each instead of a pageable copy; used only inside
  // finish_status, which leaves no copy in flight into it
  char* pin = nullptr;
  char* pin_dev = nullptr;  // its device address (kernels write it directly), or null
  size_t pin_cap = 0;
  // allocation-free call in progress (kg_ac_sweep_noalloc): every helper
  // that would allocate fails instead, and workspace names resolve to the
  // regions in `given` (the device scratch array or the caller's buffers)
  bool noalloc = false;
  std::map<std::string, std::pair<void*, size_t>> given;
  // kernel timing (HIP events on this context's stream)
  bool prof = false;
  std::map<std::string, std::pair<double, int64_t>> prof_k;  // kernel -> (ms, launches)
  std::string prof_name;  // kernel reported by the last kg_profile call
};

namespace {

int32_t set_err(kg_ctx* c, int32_t code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, x)                                                                           \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess)                                                                        \
      return set_err(ctx, KG_HIP_ERROR, std::string(#x ": ") + hipGetErrorString(e_));           \
  } while (0)

// Destroys the captured call graphs (they hold device addresses of the
// batch and of workspace buffers).
void drop_graphs(kg_ctx* c) {
  for (auto& kv : c->gexec) {
    hipGraphExecDestroy(kv.second.second);
    hipGraphDestroy(kv.second.first);
  }
  c->gexec.clear();
}

void* ws_get(kg_ctx* c, const std::string& name, size_t bytes, hipError_t* e) {
  if (c->noalloc) {
    auto it = c->given.find(name);
    if (it != c->given.end() && it->second.second >= bytes) {
      *e = hipSuccess;
      return it->second.first;
    }
    *e = hipErrorOutOfMemory;
    return nullptr;
  }
  DevBuf& b = c->ws[name];
  *e = hipSuccess;
  if (b.cap < bytes) {
    c->uploaded.erase(name);  // a new allocation holds nothing uploaded
    // a captured graph may hold the old address (not every buffer is part of
    // a graph's key): replays must not outlive it
    if (b.p) {
      drop_graphs(c);
      hipFree(b.p);  // waits for the device
    }
    b.p = nullptr;
    b.cap = 0;
    *e = hipMalloc(&b.p, std::max<size_t>(bytes, 256));
    if (*e != hipSuccess) return nullptr;
    b.cap = std::max<size_t>(bytes, 256);
  }
  return b.p;
}

void free_batch(kg_ctx* c) {
  for (auto* p : {(void*)c->d_params, (void*)c->d_ebase, (void*)c->d_ebase_t, (void*)c->d_eoff, (void*)c->d_nunk,
                  (void*)c->d_pats})
    if (p) hipFree(p);
  c->d_params = c->d_ebase = c->d_ebase_t = nullptr;
  c->d_eoff = nullptr;
  c->d_nunk = nullptr;
  c->d_pats = nullptr;
  c->d_pats_n = 0;
  for (auto& p : c->pats) {
    if (p.d_buf) hipFree(p.d_buf);
    if (p.d_lin) hipFree(p.d_lin);
  }
  c->pats.clear();
  c->pat_index.clear();
  for (auto& kv : c->inst_pat)
    if (kv.second) hipFree(kv.second);
  c->inst_pat.clear();
  c->lane_fn.clear();
  drop_graphs(c);
  c->n_inst = 0;
}

PatDev to_dev(const PatLayout& L, const int32_t* b) {
  PatDev d{};
  d.n = L.n;
  d.nnz = L.nnz;
  d.ne = L.ne;
  d.nlev = L.nlev;
  d.nblev = L.nblev;
  d.ntlev = L.ntlev;
  d.ntblev = L.ntblev;
  d.nonlinear = L.nonlinear;
  d.has_diode = L.has_diode;
  d.has_dyn = L.has_dyn;
  if (!b) return d;
  d.row_ptr = b + L.o_row_ptr;
  d.col = b + L.o_col;
  d.diag = b + L.o_diag;
  d.fprog_ptr = b + L.o_fprog_ptr;
  d.fprog = b + L.o_fprog;
  d.lvl_ptr = b + L.o_lvl_ptr;
  d.lvl_rows = b + L.o_lvl_rows;
  d.blvl_ptr = b + L.o_blvl_ptr;
  d.blvl_rows = b + L.o_blvl_rows;
  d.scon_ptr = b + L.o_scon_ptr;
  d.scon = b + L.o_scon;
  d.rcon_ptr = b + L.o_rcon_ptr;
  d.rcon = b + L.o_rcon;
  d.col_of = b + L.o_col_of;
  d.cpos = b + L.o_cpos;
  d.isbr = b + L.o_isbr;
  d.etype = b + L.o_etype;
  d.ep = b + L.o_ep;
  d.em = b + L.o_em;
  d.ec = b + L.o_ec;
  d.ut_ptr = b + L.o_ut_ptr;
  d.ut = b + L.o_ut;
  d.lt_ptr = b + L.o_lt_ptr;
  d.lt = b + L.o_lt;
  d.tlvl_ptr = b + L.o_tlvl_ptr;
  d.tlvl_rows = b + L.o_tlvl_rows;
  d.tblvl_ptr = b + L.o_tblvl_ptr;
  d.tblvl_rows = b + L.o_tblvl_rows;
  d.row_of = b + L.o_row_of;
  d.rpos = b + L.o_rpos;
  d.erp = b + L.o_erp;
  d.erm = b + L.o_erm;
  d.erc = b + L.o_erc;
  d.base = b;
  d.total = L.total;
  d.lin = nullptr;
  d.lin_words = 0;
  return d;
}

// What one analysis launch needs to know about its plans.
struct Plan {
  int32_t* d_inst_pat = nullptr;
  int nnz_max = 0, n_max = 0, ne_max = 0;
  int total_max = 0;  // longest plan buffer (int32) among the instances' patterns
  bool any_diode = false, any_nonlinear = false;
  std::vector<int32_t> group_status;  // per group
};

// s_rep: representative scale of reactive stamps used by the static pivot
// choice (alpha/h for transients, 2*pi*f for AC, unused for DC).
int32_t prepare(kg_ctx* c, AnalysisKind ak, Ordering ord, double s_rep, Plan& P) {
  if (ak == AK_DC) s_rep = 0.0;
  int ng = (int)c->group_first.size();
  bool added = false;
  P.group_status.assign(ng, KG_OK);
  std::vector<int32_t> gpi(ng);
  for (int g = 0; g < ng; g++) {
    auto key = std::make_tuple(g, (int)ak, (int)ord, s_rep);
    auto it = c->pat_index.find(key);
    int idx;
    if (it == c->pat_index.end()) {
      if (c->noalloc)
        return set_err(c, KG_INVALID_ARGUMENT,
                       "allocation-free call: no plan for this analysis yet (call the allocating entry point "
                       "once with the same batch, precision and frequency range)");
      int i0 = c->group_first[g];
      int ne = (int)(c->eoff[i0 + 1] - c->eoff[i0]);
      PatEntry pe;
      std::vector<int32_t> buf = build_pattern(c->el.data() + c->eoff[i0], ne, c->nunk[i0], ak, ord,
                                               c->params.data() + c->eoff[i0] * KG_NPARAM, s_rep, pe.lay, pe.err);
      if (pe.lay.status == KG_OK) {
        HIPCHK(c, hipMalloc(&pe.d_buf, sizeof(int32_t) * std::max<size_t>(buf.size(), 1)));
        HIPCHK(c, hipMemcpy(pe.d_buf, buf.data(), sizeof(int32_t) * buf.size(), hipMemcpyHostToDevice));
        pe.dev = to_dev(pe.lay, pe.d_buf);
        if (ak == AK_TRAN && build_lin_plan(pe.lay, buf, pe.lin)) {
          HIPCHK(c, hipMalloc(&pe.d_lin, sizeof(uint16_t) * pe.lin.size()));
          HIPCHK(c, hipMemcpy(pe.d_lin, pe.lin.data(), sizeof(uint16_t) * pe.lin.size(), hipMemcpyHostToDevice));
          pe.dev.lin = pe.d_lin;
          pe.dev.lin_words = (int32_t)pe.lin.size();
        }
        pe.host = std::move(buf);
      } else {
        PatLayout z{};
        pe.dev = to_dev(z, nullptr);
      }
      idx = (int)c->pats.size();
      c->pats.push_back(pe);
      c->pat_index[key] = idx;
      added = true;
    } else {
      idx = it->second;
    }
    gpi[g] = idx;
    const PatEntry& pe = c->pats[idx];
    P.group_status[g] = pe.lay.status;
    if (pe.lay.status == KG_OK) {
      P.nnz_max = std::max(P.nnz_max, pe.lay.nnz);
      P.n_max = std::max(P.n_max, pe.lay.n);
      P.ne_max = std::max(P.ne_max, pe.lay.ne);
      P.total_max = std::max(P.total_max, pe.lay.total);
      P.any_diode |= pe.lay.has_diode != 0;
      P.any_nonlinear |= pe.lay.nonlinear != 0;
    } else {
      c->err = pe.err;
    }
  }
  if (added || !c->d_pats) {
    if (c->noalloc) return set_err(c, KG_INVALID_ARGUMENT, "allocation-free call: plans not uploaded");
    std::vector<PatDev> hv;
    for (auto& p : c->pats) hv.push_back(p.dev);
    // kernels of this call already queued on the stream may still read the old table
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->d_pats) hipFree(c->d_pats);
    c->d_pats = nullptr;
    HIPCHK(c, hipMalloc(&c->d_pats, sizeof(PatDev) * hv.size()));
    HIPCHK(c, hipMemcpy(c->d_pats, hv.data(), sizeof(PatDev) * hv.size(), hipMemcpyHostToDevice));
    c->d_pats_n = hv.size();
  }
  auto k2 = std::make_tuple((int)ak, (int)ord, s_rep);
  if (!c->inst_pat.count(k2)) {
    if (c->noalloc) return set_err(c, KG_INVALID_ARGUMENT, "allocation-free call: plans not uploaded");
    std::vector<int32_t> tab(c->n_inst);
    for (int i = 0; i < c->n_inst; i++) tab[i] = gpi[c->group_of[i]];
    int32_t* d = nullptr;
    HIPCHK(c, hipMalloc(&d, sizeof(int32_t) * tab.size()));
    HIPCHK(c, hipMemcpy(d, tab.data(), sizeof(int32_t) * tab.size(), hipMemcpyHostToDevice));
    c->inst_pat[k2] = d;
  }
  P.d_inst_pat = c->inst_pat[k2];
  return KG_OK;
}

BatchDev batch_dev(kg_ctx* c, const Plan& P) {
  BatchDev B;
  B.pats = c->d_pats;
  B.inst_pat = P.d_inst_pat;
  B.params = c->d_params;
  B.ebase = c->d_ebase;
  B.eoff = c->d_eoff;
  B.n_inst = c->d_nunk;
  B.n_instances = c->n_inst;
  return B;
}

struct Launch {
  int tps = 64, spb = 1;
  size_t lds = 0;
  dim3 grid, block;
};

bool valid_tps(int t) { return t == 1 || t == 4 || t == 16 || t == 64 || t == 256; }

// True when some diode of the batch carries junction charge (CJO or TT of
// the opt-in diode parasitics): the run-time specialised transient kernels do
// not generate it.
bool any_junction_charge(const kg_ctx* c) {
  for (size_t r = 0; r < c->el.size(); r++)
    if (c->el[r].type == KG_D && (c->params[r * KG_NPARAM + 2] != 0.0 || c->params[r * KG_NPARAM + 6] != 0.0))
      return true;
  return false;
}

int default_tps(const kg_ctx* c, int n_max) {
  if (c->shared && n_max <= 8) return 1;
  if (n_max <= 96) return 16;
  if (n_max <= 600) return 64;
  return 256;
}

Ordering order_for(int tps) { return tps == 1 ? ORD_MINDEG : ORD_NESTED; }

// extra: bytes at the start of the block's LDS (a staged plan), once per block.
int32_t configure(kg_ctx* c, int tps, Lay lay, int64_t nsys, Launch& L, size_t extra = 0) {
  size_t sb = sys_bytes(lay);
  L.tps = tps;
  // one wave per block: small batches spread over more CUs, barriers stay cheap
  int spb = c->shared ? std::max(1, 64 / tps) : 1;
  while (spb > 1 && extra + spb * sb + 8 * (size_t)spb * tps > LDS_SHARED_BUDGET) spb /= 2;
  if (extra + spb * sb + 8 * (size_t)spb * tps > LDS_MAX)
    return set_err(c, KG_UNEXPECTED,
                   "system too large for on-chip memory (" + std::to_string(sb) + " bytes per system)");
  L.spb = spb;
  L.lds = extra + spb * sb + 8 * (size_t)spb * tps;
  L.block = dim3(tps * spb);
  L.grid = dim3((unsigned)((nsys + spb - 1) / spb));
  return KG_OK;
}

// Systems too large for LDS: one system per block of tps threads, its values
// and the block's reduction buffer in a region of a global workspace (`key`),
// launched in chunks of at most *chunk systems (lay.sys_off the first of a
// chunk) so that the workspace stays within a quarter of free device memory
// and 2048 regions (eight blocks per CU already fill the chip). The regions
// are rewritten from scratch by every system, so chunks reuse them.
constexpr int64_t KG_GWS_MAX_REGIONS = 2048;
int32_t configure_global(kg_ctx* c, int tps, Lay& lay, int64_t nsys, Launch& L, const char* key, int64_t* chunk) {
  lay.gstride = (int64_t)lay_al(sys_bytes(lay) + 8 * (size_t)tps);
  size_t free_b = 0, total_b = 0;
  HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
  const int64_t fit = (int64_t)(free_b / 4) / lay.gstride;
  if (fit < 1) return set_err(c, KG_UNEXPECTED, "system too large for device memory");
  *chunk = std::max<int64_t>(1, std::min({nsys, KG_GWS_MAX_REGIONS, fit}));
  hipError_t e;
  lay.gws = (char*)ws_get(c, key, (size_t)lay.gstride * *chunk, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (systems larger than LDS)");
  lay.sys_off = 0;
  L.tps = tps;
  L.spb = 1;
  L.lds = 0;
  L.block = dim3(tps);
  L.grid = dim3((unsigned)*chunk);
  return KG_OK;
}

// Runs launch(lay, grid) over nsys systems in chunks of `chunk` (one launch
// when the systems sit in LDS).
template <class F>
hipError_t chunked(Lay lay, int64_t nsys, int64_t chunk, const Launch& L, F&& launch) {
  if (!lay.gws) return launch(lay, L.grid);
  for (int64_t o = 0; o < nsys; o += chunk) {
    lay.sys_off = o;
    hipError_t e = launch(lay, dim3((unsigned)std::min(chunk, nsys - o)));
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Writes the status of each structurally failed topology group into its
// instances' device statuses, in stream order after the call's kernels.
int32_t mark_group_failures(kg_ctx* c, const Plan& P, int32_t* d_st) {
  bool any = false;
  for (int32_t g : P.group_status) any |= g != KG_OK;
  if (!any) return KG_OK;
  std::vector<int32_t> st(c->n_inst);
  HIPCHK(c, hipMemcpyAsync(st.data(), d_st, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->n_inst; i++)
    if (P.group_status[c->group_of[i]] != KG_OK) st[i] = P.group_status[c->group_of[i]];
  HIPCHK(c, hipMemcpyAsync(d_st, st.data(), sizeof(int32_t) * c->n_inst, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KG_OK;
}

int32_t first_failure(const std::vector<int32_t>& st) {
  for (int32_t s : st)
    if (s != KG_OK) return s;
  return KG_OK;
}

// Offsets (in doubles) of each instance's output block of `per_unknown`
// doubles per unknown.
std::vector<int64_t> out_offsets(const kg_ctx* c, int64_t per_unknown) {
  std::vector<int64_t> o(c->n_inst + 1, 0);
  for (int i = 0; i < c->n_inst; i++) o[i + 1] = o[i] + per_unknown * c->nunk[i];
  return o;
}

// Copies host data into workspace buffer `name` unless the buffer already
// holds exactly these bytes from an earlier call.
int32_t upload(kg_ctx* c, const std::string& name, const void* src, size_t bytes, void** d) {
  hipError_t e;
  *d = ws_get(c, name, bytes, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  if (c->noalloc) {  // the scratch array is shared by every call on the device: always copy
    HIPCHK(c, hipMemcpyAsync(*d, src, bytes, hipMemcpyHostToDevice, c->stream));
    return KG_OK;
  }
  auto& u = c->uploaded[name];
  if (u.first == *d && u.second.size() == bytes && std::memcmp(u.second.data(), src, bytes) == 0) return KG_OK;
  HIPCHK(c, hipMemcpyAsync(*d, src, bytes, hipMemcpyHostToDevice, c->stream));
  u.first = *d;
  u.second.assign((const char*)src, (const char*)src + bytes);
  return KG_OK;
}

int32_t upload_offsets(kg_ctx* c, const std::string& name, const std::vector<int64_t>& o, int64_t** d) {
  return upload(c, name, o.data(), sizeof(int64_t) * o.size(), (void**)d);
}

// Runs `launch` on the context stream, timing it with HIP events when
// profiling is on.
template <class F>
int32_t timed(kg_ctx* c, const char* name, F&& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  const bool prof = c->prof && !c->noalloc;
  if (prof) {
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipEventRecord(a, c->stream));
  }
  hipError_t e = launch();
  if (e != hipSuccess) {
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    return set_err(c, KG_HIP_ERROR, std::string(name) + ": " + hipGetErrorString(e));
  }
  if (prof) {
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, a, b));
    auto& k = c->prof_k[name];
    k.first += ms;
    k.second += 1;
    hipEventDestroy(a);
    hipEventDestroy(b);
  }
  return KG_OK;
}

// ---------------------------------------------------------------- autotuner
// Picks the work mapping (threads per system and, with it, the ordering: a
// minimum-degree order for one thread per system, nested dissection for
// level-scheduled groups) by timing every feasible candidate on the actual
// workload once, then caches the winner per workload key.
void tune_load(kg_ctx* c) {
  const char* f = std::getenv("KRETS_GPU_TUNE_FILE");
  if (!f || !*f) return;
  c->tune_file = f;
  std::ifstream in(f);
  std::string line;
  while (std::getline(in, line)) {  // lines: "key" : tps
    size_t a = line.find('"'), b = line.find('"', a + 1), col = line.find(':', b);
    if (a == std::string::npos || b == std::string::npos || col == std::string::npos) continue;
    c->tune[line.substr(a + 1, b - a - 1)] = std::atoi(line.c_str() + col + 1);
  }
}

void tune_save(kg_ctx* c) {
  if (c->tune_file.empty()) return;
  std::ofstream out(c->tune_file);
  out << "{\n";
  size_t k = 0;
  for (auto& kv : c->tune) out << "  \"" << kv.first << "\": " << kv.second << (++k < c->tune.size() ? ",\n" : "\n");
  out << "}\n";
}

std::string tune_key(const kg_ctx* c, const char* what, int precision, const Plan& P, int64_t nsys) {
  std::ostringstream k;
  int lg = 0;
  while ((int64_t(1) << (lg + 1)) <= nsys) lg++;
  k << what << "|p" << precision << "|n" << P.n_max << "|nnz" << P.nnz_max << "|ne" << P.ne_max << "|s2^" << lg
    << (c->shared ? "|shared" : "|uneven");
  return k.str();
}

// trial(tps, &ms) runs the workload with that mapping; returns a kg_status.
int32_t autotune(kg_ctx* c, const std::string& key, const std::vector<int>& cands,
                 const std::function<int32_t(int, float*)>& trial, int* best) {
  auto it = c->tune.find(key);
  if (it != c->tune.end()) {
    *best = it->second;
    return KG_OK;
  }
  c->last_trials.clear();
  float best_ms = 0.f;
  *best = 0;
  for (int tps : cands) {
    float ms = 0.f;
    int32_t r = trial(tps, &ms);
    if (r == KG_HIP_ERROR) return r;
    if (r != KG_OK) continue;  // infeasible mapping (on-chip memory)
    c->last_trials.push_back({tps, ms});
    if (*best == 0 || ms < best_ms) {
      best_ms = ms;
      *best = tps;
    }
  }
  if (*best == 0) return set_err(c, KG_UNEXPECTED, "no feasible work mapping for this workload");
  c->tune[key] = *best;
  tune_save(c);
  c->err.clear();
  return KG_OK;
}

// Times one launch with HIP events (used by the autotuner).
template <class F>
int32_t time_launch(kg_ctx* c, F&& launch, float* ms) {
  hipEvent_t a, b;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  HIPCHK(c, hipEventRecord(a, c->stream));
  hipError_t e = launch();
  if (e != hipSuccess) {
    hipEventDestroy(a);
    hipEventDestroy(b);
    return set_err(c, KG_HIP_ERROR, hipGetErrorString(e));
  }
  HIPCHK(c, hipEventRecord(b, c->stream));
  HIPCHK(c, hipEventSynchronize(b));
  HIPCHK(c, hipEventElapsedTime(ms, a, b));
  hipEventDestroy(a);
  hipEventDestroy(b);
  return KG_OK;
}

NewtonArgs newton_args(const kg_solver_config* cfg) {
  return NewtonArgs{cfg->reltol, cfg->iabstol, cfg->vabstol, cfg->maxiter, cfg->compat};
}

// Runs OP (npts = 1) or a DC sweep; result stays on the device in the
// workspace buffer "newton_out"; status in "status" (int32 per instance).
// A launch that can run eagerly (timed with events) or be captured.
struct Deferred {
  const char* name;  // nullptr: not timed
  std::function<hipError_t()> launch;
};

int32_t run_dc(kg_ctx* c, const kg_solver_config* cfg, int sweep_e, double start, double step, int npts,
               double** d_out, int64_t** d_off, int32_t** d_status, int32_t** d_iters,
               std::vector<Deferred>* deferred = nullptr) {
  Plan P;
  int tps0 = cfg->threads_per_system > 0 ? cfg->threads_per_system : 0;
  // prepare with a provisional ordering to learn sizes
  Plan P0;
  int32_t r = prepare(c, AK_DC, ORD_NESTED, 0.0, P0);
  if (r) return r;
  // linear circuits of a shared topology: one solve per point, in the lane
  // kernel specialised to the topology at run time (2001), unless another
  // mapping is asked for or the topology does not fit it
  if (c->shared && P0.group_status[0] == KG_OK && cfg->maxiter >= 1 && (tps0 == 0 || tps0 == 2001)) {
    Plan PL;
    if ((r = prepare(c, AK_DC, ORD_MINDEG, 0.0, PL))) return r;
    const int pidx = c->pat_index[std::make_tuple(0, (int)AK_DC, (int)ORD_MINDEG, 0.0)];
    const auto key = std::make_tuple(pidx, 200, 0);
    auto it = c->lane_fn.find(key);
    JitFn fn;
    std::string why;
    if (it != c->lane_fn.end()) {
      fn = it->second;
    } else {
      if (c->noalloc) return set_err(c, KG_INVALID_ARGUMENT, "allocation-free call: kernel not compiled yet");
      const PatEntry& pe = c->pats[pidx];
      std::string src, log;
      if (jit_dc_source(pe.lay, pe.host.data(), src, why)) {
        if (jit_load(c->device, src, jit_dc_name(), fn, log) != hipSuccess)
          return set_err(c, KG_UNEXPECTED, "lane kernel compilation failed: " + log);
      }
      c->lane_fn[key] = fn;  // an empty entry records a topology the generator declines
    }
    if (fn.fn) {
      auto off = out_offsets(c, npts);
      if ((r = upload_offsets(c, "newton_off", off, d_off))) return r;
      hipError_t e;
      *d_out = (double*)ws_get(c, "newton_out", sizeof(double) * std::max<int64_t>(off.back(), 1), &e);
      if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
      *d_status = (int32_t*)ws_get(c, "status", sizeof(int32_t) * c->n_inst, &e);
      *d_iters = (int32_t*)ws_get(c, "newton_iters", sizeof(int32_t) * (size_t)c->n_inst * npts, &e);
      if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
      const double* prm = c->d_params;
      const double* eb = c->d_ebase;
      int32_t *st = *d_status, *its = *d_iters;
      double* o = *d_out;
      int64_t* oo = *d_off;
      hipStream_t strm = c->stream;
      const int ni = c->n_inst;
      const hipFunction_t f = fn.fn;
      const NewtonArgs na = newton_args(cfg);
      std::vector<Deferred> steps{{jit_dc_name(), [=]() mutable {
                                     int a_ni = ni, a_se = sweep_e, a_np = npts, a_cp = na.compat, a_mi = na.maxiter;
                                     double a_st = start, a_sp = step, a_rt = na.reltol, a_it = na.iabstol,
                                            a_vt = na.vabstol;
                                     void* args[] = {&prm, &eb,  &a_ni, &a_se, &a_st, &a_sp, &a_np, &o,   &oo,
                                                     &st,  &its, &a_cp, &a_rt, &a_it, &a_vt, &a_mi};
                                     return hipModuleLaunchKernel(f, (unsigned)((ni + 63) / 64), 1, 1, 64, 1, 1, 0, strm,
                                                                  args, nullptr);
                                   }}};
      if (deferred) {
        deferred->insert(deferred->end(), steps.begin(), steps.end());
        return KG_OK;
      }
      return timed(c, steps[0].name, steps[0].launch);
    }
    if (tps0 == 2001) return set_err(c, KG_UNEXPECTED, "lane kernel: " + why);
  }
  // a few large systems (a single operating point of a long ladder): a whole
  // CU per system, 512 threads over each elimination level (1024 halves the
  // registers per thread to 128, and the kernel spills there)
  int tps = tps0 ? tps0 : (P0.n_max > 600 && c->n_inst <= 64 ? 512 : default_tps(c, P0.n_max));
  if (!valid_tps(tps) && tps != 512 && tps != 1024)
    return set_err(c, KG_INVALID_ARGUMENT, "threads_per_system must be 1, 4, 16, 64, 256, 512 or 1024");
  r = prepare(c, AK_DC, order_for(tps), 0.0, P);
  if (r) return r;
  Lay lay{P.nnz_max, P.n_max, P.ne_max, 1, 0, P.any_diode ? 1 : 0, 8, 8, 0, 0};
  Launch L;
  hipError_t e;
  int64_t chunk = c->n_inst;
  if ((r = configure(c, tps, lay, c->n_inst, L))) {
    if (tps < 64) return r;
    c->err.clear();  // too large for LDS: the values go to a global workspace
    if ((r = configure_global(c, tps, lay, c->n_inst, L, "newton_gws", &chunk))) return r;
  } else if (c->shared && tps >= 256) {
    // one large system per block: its factor and solve structure goes to LDS too when it fits
    const PatEntry& pe = c->pats[c->pat_index[std::make_tuple(0, (int)AK_DC, (int)order_for(tps), 0.0)]];
    const size_t add = sizeof(int32_t) * (size_t)pe.lay.o_col;
    if (pe.lay.status == KG_OK && L.lds + add <= LDS_MAX) {
      lay.stage_plan = pe.lay.o_col;
      L.lds += add;
    }
  }
  auto off = out_offsets(c, npts);
  r = upload_offsets(c, "newton_off", off, d_off);
  if (r) return r;
  *d_out = (double*)ws_get(c, "newton_out", sizeof(double) * std::max<int64_t>(off.back(), 1), &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  *d_status = (int32_t*)ws_get(c, "status", sizeof(int32_t) * c->n_inst, &e);
  *d_iters = (int32_t*)ws_get(c, "newton_iters", sizeof(int32_t) * (size_t)c->n_inst * npts, &e);
  BatchDev B = batch_dev(c, P);
  NewtonArgs na = newton_args(cfg);
  int32_t* st = *d_status;
  int32_t* its = *d_iters;
  double* o = *d_out;
  int64_t* oo = *d_off;
  hipStream_t strm = c->stream;
  int ni = c->n_inst;
  std::vector<Deferred> steps{
      {nullptr, [=] { return run_fill_i32(st, KG_OK, ni, strm); }},
      {"k_newton_dc", [=] {
         return chunked(lay, ni, chunk, L, [&](Lay ly, dim3 g) {
           return run_newton(tps, g, L.block, L.lds, strm, B, 0, na, sweep_e, start, step, npts, o, oo, st, its, ly);
         });
       }}};
  // instances whose topology has no plan fail as krets' sp_lu would
  bool any_bad = false;
  for (auto s : P.group_status) any_bad |= s != KG_OK;
  if (deferred && !any_bad) {
    deferred->insert(deferred->end(), steps.begin(), steps.end());
    return KG_OK;
  }
  for (auto& d : steps) {
    int32_t rr = d.name ? timed(c, d.name, d.launch) : KG_OK;
    if (!d.name) HIPCHK(c, d.launch());
    if (rr) return rr;
  }
  if (any_bad) {
    std::vector<int32_t> st(c->n_inst);
    HIPCHK(c, hipMemcpyAsync(st.data(), *d_status, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->n_inst; i++)
      if (P.group_status[c->group_of[i]] != KG_OK) st[i] = P.group_status[c->group_of[i]];
    HIPCHK(c, hipMemcpyAsync(*d_status, st.data(), sizeof(int32_t) * c->n_inst, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // st is a local host buffer
  }
  return KG_OK;
}

// LDS bytes the linear-transient kernel (k_tran_lin) needs for instance i,
// 0 if its topology has no compact plan; instances without a valid plan
// return 1 (they are never integrated, their status is already an error).
size_t lin_need_one(kg_ctx* c, double s_tr, int i) {
  auto it = c->pat_index.find(std::make_tuple(c->group_of[i], (int)AK_TRAN, (int)ORD_NESTED, s_tr));
  if (it == c->pat_index.end()) return 0;
  const PatEntry& pe = c->pats[it->second];
  if (pe.lay.status != KG_OK) return 1;
  if (pe.lin.empty()) return 0;
  LinHdr h;
  std::memcpy(&h, pe.lin.data(), sizeof(h));
  return lin_bytes(h);
}

// Most independent sources of any instance's compact plan.
int lin_nsrc_max(kg_ctx* c, double s_tr) {
  int m = 0;
  for (int i = 0; i < c->n_inst; i++) {
    auto it = c->pat_index.find(std::make_tuple(c->group_of[i], (int)AK_TRAN, (int)ORD_NESTED, s_tr));
    if (it == c->pat_index.end() || c->pats[it->second].lin.empty()) continue;
    LinHdr h;
    std::memcpy(&h, c->pats[it->second].lin.data(), sizeof(h));
    m = std::max(m, h.nsrc);
  }
  return m;
}

// Source-value table of k_tran_lin (n_instances * nsrc_max * (nsteps + 1)).
double* lin_srcv(kg_ctx* c, double s_tr, int nsteps, int* nsrc_max, int32_t* r) {
  *nsrc_max = lin_nsrc_max(c, s_tr);
  hipError_t e;
  double* p = (double*)ws_get(c, "tran_srcv",
                              sizeof(double) * std::max<size_t>(1, (size_t)c->n_inst * *nsrc_max * (nsteps + 1)), &e);
  *r = e == hipSuccess ? KG_OK : set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  return p;
}

// LDS bytes of the LDS-resident adjoint for the (shared) topology; 0 if it has
// no compact plan.
size_t lin_adj_need(kg_ctx* c, double s_tr) {
  auto it = c->pat_index.find(std::make_tuple(c->group_of[0], (int)AK_TRAN, (int)ORD_NESTED, s_tr));
  if (it == c->pat_index.end() || c->pats[it->second].lin.empty()) return 0;
  LinHdr h;
  std::memcpy(&h, c->pats[it->second].lin.data(), sizeof(h));
  return lin_adj_bytes(h);
}

// Source-value table of the adjoint kernels (n_instances * nsrc_max * (nsteps + 1)).
double* adj_srcv(kg_ctx* c, double s_tr, int nsteps, int* nsrc_max, int32_t* r) {
  *nsrc_max = lin_nsrc_max(c, s_tr);
  hipError_t e;
  double* p = (double*)ws_get(c, "adj_srcv",
                              sizeof(double) * std::max<size_t>(1, (size_t)c->n_inst * *nsrc_max * (nsteps + 1)), &e);
  *r = e == hipSuccess ? KG_OK : set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  return p;
}

// Largest need over all instances, 0 if any instance has no compact plan.
size_t lin_need_all(kg_ctx* c, double s_tr) {
  size_t m = 0;
  for (int i = 0; i < c->n_inst; i++) {
    size_t b = lin_need_one(c, s_tr, i);
    if (b == 0) return 0;
    m = std::max(m, b);
  }
  return m;
}

// Work-mapping codes of k_tran_lin: 3000 + threads per block (one system per block).
const std::vector<int> LIN_CANDS = {3064, 3128, 3256, 3512, 4024};
bool is_lin_cand(int t) { return t == 3064 || t == 3128 || t == 3256 || t == 3512 || t == 4024; }
// Rows per thread of the LDS-resident adjoint kernels (64, 256, 512 threads;
// 512 threads at most 2 rows); 0 if none fits.
int adj_lin_rows(int tps, int n_max) {
  if (tps != 64 && tps != 256 && tps != 512) return 0;
  for (int r : {1, 2, 4})
    if (n_max <= r * tps && (tps < 512 || r <= 2)) return r;
  return 0;
}

// Rows per thread of k_tran_lin for systems of up to n_max unknowns; 0 if none fits.
int lin_rows(int tps, int n_max) {
  for (int r : {1, 2, 4})
    if (n_max <= r * tps && (tps < 1024 || r <= 2)) return r;
  return 0;
}

// Uneven transient batches: every block of one launch reserves LDS for the
// largest instance, so a launch over 8..512-section ladders runs one block
// per CU. Instances are split into size classes by how many of their blocks
// fit one CU's LDS; each class gets its own launch, LDS footprint and
// autotuned mapping, on its own stream so that classes overlap.
int32_t run_tran_buckets(kg_ctx* c, const NewtonArgs& na, const TranArgs& ta, bool mixed, double s_tr,
                         int precision, const double* d_xop, const int64_t* d_xoff, double* d_out,
                         const int64_t* d_off, int32_t* d_st, const int32_t* d_st0, int32_t* d_iters,
                         double* d_err) {
  Plan P;
  int32_t r = prepare(c, AK_TRAN, ORD_NESTED, s_tr, P);
  if (r) return r;
  const size_t red = 8 * 256;
  // FP64 linear batches are classed by the LDS footprint of k_tran_lin
  const bool lin_ok = !mixed && lin_need_all(c, s_tr) > 0;
  std::map<int, std::vector<int32_t>> cls;  // blocks per CU -> instances
  std::map<int, Lay> clay;
  std::map<int, int> cplan;      // longest plan per class
  std::map<int, size_t> clin;    // largest k_tran_lin footprint per class
  for (int i = 0; i < c->n_inst; i++) {
    const PatEntry& pe = c->pats[c->pat_index[std::make_tuple(c->group_of[i], (int)AK_TRAN, (int)ORD_NESTED, s_tr)]];
    if (pe.lay.status != KG_OK) continue;  // structurally singular: status set by the caller
    Lay l{pe.lay.nnz, pe.lay.n, pe.lay.ne, 1, 1, 1, mixed ? 4 : 8, 8, mixed ? 1 : 0, 0};
    // class -1: systems larger than LDS, their values in a global workspace
    const bool big = sys_bytes(l) + red > LDS_MAX;
    const size_t lb = lin_ok && !big ? lin_need_one(c, s_tr, i) : 0;
    int per_cu = big ? -1 : (int)std::min<size_t>(8, LDS_MAX / (lin_ok ? lb : sys_bytes(l) + red));
    cls[per_cu].push_back(i);
    Lay& b = clay.emplace(per_cu, l).first->second;
    b.nnz_max = std::max(b.nnz_max, l.nnz_max);
    b.n_max = std::max(b.n_max, l.n_max);
    b.ne_max = std::max(b.ne_max, l.ne_max);
    cplan[per_cu] = std::max(cplan[per_cu], pe.lay.total);
    clin[per_cu] = std::max(clin[per_cu], lb);
  }
  struct Bucket {
    Lay lay;
    Launch L;
    int tps = 0;  // 1000 + t: plan staged in LDS; 3000 + t: k_tran_lin
    int32_t* list = nullptr;
    int32_t n = 0;
    int plan_ints = 0;
    size_t lin = 0;
    Launch gL;           // class -1: launch shape for 256 threads (workspace regions sized for it)
    int64_t gchunk = 0;  // class -1: systems per launch
  };
  std::vector<Bucket> bk;
  for (auto& kv : cls) {
    Bucket b;
    b.lay = clay[kv.first];
    b.n = (int32_t)kv.second.size();
    b.plan_ints = cplan[kv.first];
    b.lin = clin[kv.first];
    hipError_t e;
    b.list = (int32_t*)ws_get(c, "tran_bucket" + std::to_string(kv.first), sizeof(int32_t) * kv.second.size(), &e);
    if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
    HIPCHK(c, hipMemcpyAsync(b.list, kv.second.data(), sizeof(int32_t) * kv.second.size(), hipMemcpyHostToDevice,
                             c->stream));
    if (kv.first < 0 && (r = configure_global(c, 256, b.lay, b.n, b.gL, "tran_gws", &b.gchunk))) return r;
    bk.push_back(b);
  }
  std::sort(bk.begin(), bk.end(), [](const Bucket& a, const Bucket& b) { return a.lay.nnz_max > b.lay.nnz_max; });
  BatchDev B0 = batch_dev(c, P);
  int nsrc_max = 0;
  double* srcv = nullptr;
  if (lin_ok) {
    srcv = lin_srcv(c, s_tr, ta.nsteps, &nsrc_max, &r);
    if (r) return r;
  }
  auto launch = [&](const Bucket& b, int cand, hipStream_t st) -> hipError_t {
    if (is_lin_cand(cand)) {
      BatchDev B = B0;
      B.inst_list = b.list;
      B.n_list = b.n;
      const int rows = lin_rows(cand - 3000, b.lay.n_max);
      if (rows == 0) return hipErrorInvalidConfiguration;
      return run_tran_lin(cand - 3000, rows, dim3((unsigned)b.n), b.lin, st, B, 0, ta, d_xop, d_xoff, d_out, d_off,
                          d_st, d_iters, d_err, srcv, nsrc_max);
    }
    Launch L;
    int tps = cand % 1000, pi = cand >= 1000 ? b.plan_ints : 0;
    if (b.lay.gws) {
      if (pi || tps > 256) return hipErrorInvalidConfiguration;
      L = b.gL;
      L.tps = tps;
      L.block = dim3(tps);
    } else if (configure(c, tps, b.lay, b.n, L, pi ? lay_al(4 * (size_t)pi) : 0) != KG_OK) {
      return hipErrorInvalidConfiguration;
    }
    BatchDev B = B0;
    B.inst_list = b.list;
    B.n_list = b.n;
    return chunked(b.lay, b.n, b.gchunk, L, [&](Lay ly, dim3 g) {
      return run_tran(tps, mixed ? 1 : 0, g, L.block, L.lds, st, B, 0, na, ta, d_xop, d_xoff, d_out, d_off, d_st,
                      d_iters, d_err, ly, pi);
    });
  };
  // mapping per size class, timed on that class alone
  for (auto& b : bk) {
    Plan Pb = P;
    Pb.n_max = b.lay.n_max;
    Pb.nnz_max = b.lay.nnz_max;
    Pb.ne_max = b.lay.ne_max;
    std::vector<int> cands = {64, 256};
    if (!b.lay.gws) cands.insert(cands.end(), {1064, 1256});
    if (lin_ok && !b.lay.gws && b.lin <= LDS_MAX) cands.insert(cands.end(), LIN_CANDS.begin(), LIN_CANDS.end());
    r = autotune(c, tune_key(c, lin_ok ? "tranl" : "tran", precision, Pb, b.n), cands,
                 [&](int t, float* ms) -> int32_t {
      if (b.lay.gws) {
        if (t >= 1000) return set_err(c, KG_UNEXPECTED, "systems larger than LDS: plan not staged");
      } else if (!is_lin_cand(t)) {
        Launch L;
        int32_t rr = configure(c, t % 1000, b.lay, b.n, L, t >= 1000 ? lay_al(4 * (size_t)b.plan_ints) : 0);
        if (rr) return rr;
      } else if (lin_rows(t - 3000, b.lay.n_max) == 0) {
        return set_err(c, KG_UNEXPECTED, "more than 4 rows per thread");
      }
      HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToDevice, c->stream));
      return time_launch(c, [&] { return launch(b, t, c->stream); }, ms);
    }, &b.tps);
    if (r) return r;
  }
  HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToDevice, c->stream));
  while (c->aux.size() + 1 < bk.size()) {
    hipStream_t s;
    HIPCHK(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    c->aux.push_back(s);
  }
  bool all_lin = true;
  for (auto& b : bk) all_lin = all_lin && is_lin_cand(b.tps);
  const char* kname = mixed ? "k_tran<float,double>" : all_lin ? "k_tran_lin" : "k_tran<double,double>";
  return timed(c, kname, [&]() -> hipError_t {
    hipEvent_t fork;
    hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    std::vector<hipEvent_t> done;
    if ((e = hipEventRecord(fork, c->stream)) != hipSuccess) return e;
    for (size_t k = 0; k < bk.size(); k++) {
      hipStream_t st = k == 0 ? c->stream : c->aux[k - 1];
      if (k > 0 && (e = hipStreamWaitEvent(st, fork, 0)) != hipSuccess) return e;
      if ((e = launch(bk[k], bk[k].tps, st)) != hipSuccess) return e;
      if (k > 0) {
        hipEvent_t d;
        if ((e = hipEventCreateWithFlags(&d, hipEventDisableTiming)) != hipSuccess) return e;
        if ((e = hipEventRecord(d, st)) != hipSuccess) return e;
        done.push_back(d);
      }
    }
    for (hipEvent_t d : done) {
      if ((e = hipStreamWaitEvent(c->stream, d, 0)) != hipSuccess) return e;
      hipEventDestroy(d);
    }
    hipEventDestroy(fork);
    return hipSuccess;
  });
}

// A result handed back to the caller: bytes from the device array src into
// dst. big: one of the call's large outputs, whose memory `mem` names; the
// small per-instance ones (iterations, error figures) may be either (the
// pointer decides).
struct Out {
  void* dst;
  const void* src;
  size_t bytes;
  bool big;
};
int32_t finish_status(kg_ctx* c, const int32_t* d_status, int32_t* status_out, int mem,
                      std::initializer_list<Out> outs = {});
int32_t run_adj_op_term(kg_ctx* c, const Plan& P, double* d_xb, double* d_g, int ne, int32_t* d_st);

// Backward pass with the LDS-resident adjoint kernel (FP64), then the
// operating-point term.
int32_t adjoint_backward_lin(kg_ctx* c, double* grad, int32_t* status, int32_t mem) {
  const AdjArgs& aa = c->adj.aa;
  const double s_tr = aa.alpha / aa.h;
  Plan P;
  int32_t r;
  if ((r = prepare(c, AK_TRAN, ORD_NESTED, s_tr, P))) return r;
  const int tps = c->adj.tps - 3000, rows = adj_lin_rows(tps, P.n_max);
  const size_t ladj = lin_adj_need(c, s_tr);
  if (rows == 0 || ladj == 0 || ladj > LDS_MAX) return set_err(c, KG_UNEXPECTED, "adjoint mapping no longer fits");
  hipError_t e;
  size_t ni = c->n_inst;
  int ne = (int)(c->eoff[1] - c->eoff[0]);
  int64_t seg_stride = (int64_t)(aa.K + 1) * P.n_max;
  double* d_ck = (double*)ws_get(c, "adj_ckpt", 8, &e);
  double* d_c = (double*)ws_get(c, "adj_cvec", 8, &e);
  double* d_seg = (double*)ws_get(c, "adj_seg", sizeof(double) * ni * seg_stride, &e);
  double* d_g = (double*)ws_get(c, "adj_grad", sizeof(double) * ni * ne, &e);
  double* d_xb = (double*)ws_get(c, "adj_xbar0", sizeof(double) * ni * P.n_max, &e);
  int64_t* d_xoff = (int64_t*)ws_get(c, "adj_xoff", 8, &e);
  int32_t* d_st = (int32_t*)ws_get(c, "adj_status", sizeof(int32_t) * ni, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (adjoint)");
  int nsrc_max = 0;
  double* srcv = adj_srcv(c, s_tr, aa.nsteps, &nsrc_max, &r);
  if (r) return r;
  HIPCHK(c, run_fill_i32(d_st, KG_OK, (int64_t)ni, c->stream));
  TranArgs ta{aa.h, aa.alpha, aa.nsteps, aa.method, 0};
  BatchDev B = batch_dev(c, P);
  const int64_t ck_stride = (int64_t)c->adj.ncp * c->adj.cpw;
  // Every state for the backward pass, so that it recomputes nothing: with a
  // checkpoint per step they are the checkpoints; otherwise, when the whole
  // trajectory fits in a quarter of the free device memory, all segments
  // are recomputed in parallel from their checkpoints first (k_adj_lin_rec),
  // instead of one after another inside the backward sweep.
  const double* xs = nullptr;
  int64_t xs_inst = 0, xs_step = 0;
  if (aa.K == 1) {
    xs = d_ck;
    xs_inst = ck_stride;
    xs_step = c->adj.cpw;
  } else {
    const int64_t traj_stride = (int64_t)(aa.nsteps + 1) * P.n_max;
    const size_t traj_bytes = sizeof(double) * ni * (size_t)traj_stride;
    size_t free_b = 0, total_b = 0;
    const size_t lfwd = lin_need_one(c, s_tr, 0);
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && traj_bytes <= free_b / 4 && lfwd > 1 && lfwd <= LDS_MAX) {
      double* d_traj = (double*)ws_get(c, "adj_traj", traj_bytes, &e);
      if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (adjoint trajectory)");
      const int64_t nseg = (aa.nsteps + aa.K - 1) / aa.K;
      r = timed(c, "k_adj_lin_rec", [&] {
        return run_adj_lin_rec(tps, rows, dim3((unsigned)(ni * nseg)), lfwd, c->stream, B, 0, ta, aa.K, d_ck,
                               ck_stride, c->adj.cpw, d_traj, traj_stride, d_st, srcv, nsrc_max);
      });
      if (r) return r;
      xs = d_traj;
      xs_inst = traj_stride;
      xs_step = P.n_max;
    } else {
      (void)hipGetLastError();
    }
  }
  r = timed(c, "k_adj_lin_bwd", [&] {
    return run_adj_lin_bwd(tps, rows, dim3((unsigned)ni), ladj, c->stream, B, 0, ta, aa.K, d_c, d_ck, ck_stride,
                           c->adj.cpw, d_seg, seg_stride, d_g, ne, d_xb, d_xoff, d_st, srcv, nsrc_max, xs, xs_inst,
                           xs_step);
  });
  if (r) return r;
  if ((r = run_adj_op_term(c, P, d_xb, d_g, ne, d_st))) return r;
  return finish_status(c, d_st, status, mem, {{grad, d_g, sizeof(double) * ni * ne, true}});
}

// Operating-point part of the adjoint (DC plan, FP64): grad += dJ/dx_0 terms.
int32_t run_adj_op_term(kg_ctx* c, const Plan& P, double* d_xb, double* d_g, int ne, int32_t* d_st) {
  hipError_t e;
  double* d_xop = (double*)ws_get(c, "adj_xop", 8, &e);
  int64_t* d_xoff = (int64_t*)ws_get(c, "adj_xoff", 8, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (adjoint)");
  Plan PD;
  int tpd = default_tps(c, P.n_max);
  int32_t r;
  if ((r = prepare(c, AK_DC, order_for(tpd), 0.0, PD))) return r;
  Lay ld{PD.nnz_max, PD.n_max, PD.ne_max, 0, 0, 0, 8, 8, 0, 0};
  Launch LD;
  if ((r = configure(c, tpd, ld, c->n_inst, LD))) return r;
  BatchDev BD = batch_dev(c, PD);
  return timed(c, "k_adj_op", [&] {
    return run_adj_op(tpd, LD.grid, LD.block, LD.lds, c->stream, BD, 0, d_xb, d_xop, d_xoff, d_g, ne, d_st, ld);
  });
}

// True when p is device memory of the context's GPU (host memory, pageable
// or pinned, is not).
bool on_device(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable host memory is unknown to the runtime
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// The end of every analysis call: copies the statuses and the call's results
// to the caller, synchronises the stream and returns the first failure.
// Device destinations get device copies. Host destinations up to
// KG_PIN_STAGE bytes in all are written by one copy kernel into the pinned
// staging area (with a device address; else by DMA copies into it) and
// copied out after the synchronisation: a copy into pageable memory costs
// more than the kernel. Larger host results are copied directly. Every copy
// enqueued here has completed when it returns.
constexpr size_t KG_PIN_STAGE = 4u << 20;
int32_t finish_status(kg_ctx* c, const int32_t* d_status, int32_t* status_out, int mem,
                      std::initializer_list<Out> outs) {
  const size_t sb = sizeof(int32_t) * c->n_inst;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  std::vector<Out> staged;
  std::vector<size_t> at;
  size_t need = al(sb);
  hipError_t e = hipSuccess;
  for (const Out& o : outs) {
    if (!o.dst || o.dst == o.src || o.bytes == 0) continue;
    const bool dev = o.big ? mem == KG_MEM_DEVICE : on_device(o.dst);
    if (dev) {
      if (e == hipSuccess) e = hipMemcpyAsync(o.dst, o.src, o.bytes, hipMemcpyDeviceToDevice, c->stream);
    } else if (need + al(o.bytes) <= KG_PIN_STAGE && staged.size() + 1 < (size_t)KG_COPY_SEGS) {
      staged.push_back(o);
      at.push_back(need);
      need += al(o.bytes);
    } else if (e == hipSuccess) {
      e = hipMemcpyAsync(o.dst, o.src, o.bytes, hipMemcpyDeviceToHost, c->stream);
    }
  }
  const bool shost = status_out && !on_device(status_out);
  if (need > c->pin_cap && c->noalloc) {
    // no staging area to read the statuses through: they stay in the
    // caller's device array, which the kernels wrote directly
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e != hipSuccess || es != hipSuccess)
      return set_err(c, KG_HIP_ERROR, std::string("result copy: ") + hipGetErrorString(e != hipSuccess ? e : es));
    return KG_OK;
  }
  if (need > c->pin_cap) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->pin) (void)hipHostFree(c->pin);
    c->pin = c->pin_dev = nullptr;
    c->pin_cap = 0;
    const size_t cap = std::max<size_t>(2 * need, 1 << 16);
    HIPCHK(c, hipHostMalloc((void**)&c->pin, cap, hipHostMallocDefault));
    c->pin_cap = cap;
    // with a device address the results are published by one small kernel
    // (DMA copies cost several times its launch)
    if (hipHostGetDevicePointer((void**)&c->pin_dev, c->pin, 0) != hipSuccess) {
      c->pin_dev = nullptr;
      (void)hipGetLastError();
    }
  }
  if (e == hipSuccess && c->pin_dev) {
    CopySegs cs{};
    cs.n = 0;
    auto seg = [&](const void* src, size_t off, size_t bytes) {
      cs.src[cs.n] = (const uint32_t*)src;
      cs.dst[cs.n] = (uint32_t*)(c->pin_dev + off);
      cs.start[cs.n + 1] = cs.start[cs.n] + (int64_t)(bytes / 4);
      cs.n++;
    };
    seg(d_status, 0, sb);
    for (size_t k = 0; k < staged.size(); k++) seg(staged[k].src, at[k], staged[k].bytes);
    e = run_copy_segments(cs, c->stream);
  } else if (e == hipSuccess) {
    e = hipMemcpyAsync(c->pin, d_status, sb, hipMemcpyDeviceToHost, c->stream);
    for (size_t k = 0; k < staged.size() && e == hipSuccess; k++)
      e = hipMemcpyAsync(c->pin + at[k], staged[k].src, staged[k].bytes, hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess && status_out && !shost && status_out != d_status)
    e = hipMemcpyAsync(status_out, d_status, sb, hipMemcpyDeviceToDevice, c->stream);
  const hipError_t es = hipStreamSynchronize(c->stream);
  if (e != hipSuccess || es != hipSuccess)
    return set_err(c, KG_HIP_ERROR, std::string("result copy: ") + hipGetErrorString(e != hipSuccess ? e : es));
  std::vector<int32_t> st((const int32_t*)c->pin, (const int32_t*)c->pin + c->n_inst);
  if (shost) std::memcpy(status_out, c->pin, sb);
  for (size_t k = 0; k < staged.size(); k++) std::memcpy(staged[k].dst, c->pin + at[k], staged[k].bytes);
  int32_t f = first_failure(st);
  if (f != KG_OK && c->err.empty()) c->err = "at least one instance failed";
  return f;
}

}  // namespace

extern "C" {

void kg_default_config(kg_solver_config* cfg) {
  cfg->reltol = 1e-3;  // config.rs:25-36
  cfg->iabstol = 1e-12;
  cfg->vabstol = 1e-6;
  cfg->maxiter = 300;
  cfg->compat = 1;
  cfg->threads_per_system = 0;
  cfg->refine_iters = 0;
}

int32_t kg_create(kg_ctx** out, int32_t device) {
  if (!out) return KG_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return KG_HIP_ERROR;
  if (device < 0 || device >= n) return KG_INVALID_ARGUMENT;
  auto* c = new kg_ctx();
  c->device = device;
  tune_load(c);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return KG_HIP_ERROR;
  }
  scratch_init(device);
  *out = c;
  return KG_OK;
}

int32_t kg_destroy(kg_ctx* c) {
  if (!c) return KG_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  free_batch(c);
  for (auto& kv : c->ws)
    if (kv.second.p) hipFree(kv.second.p);
  for (hipStream_t s : c->aux) hipStreamDestroy(s);
  hipStreamDestroy(c->stream);
  if (c->pin) hipHostFree(c->pin);
  delete c;
  return KG_OK;
}

const char* kg_last_error(const kg_ctx* c) { return c ? c->err.c_str() : "no context"; }

int32_t kg_load_batch(kg_ctx* c, const kg_batch_desc* b) {
  if (!c || !b || b->n_instances <= 0 || !b->elements || !b->params || !b->n_unknowns)
    return set_err(c, KG_INVALID_ARGUMENT, "invalid batch descriptor");
  c->err.clear();
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_batch(c);
  int ni = b->n_instances;
  c->n_inst = ni;
  c->shared = b->shared_topology != 0;
  c->eoff.assign(ni + 1, 0);
  c->nunk.assign(ni, 0);
  if (c->shared) {
    int ne = b->n_elements;
    c->el.assign(b->elements, b->elements + ne);
    for (int i = 0; i <= ni; i++) c->eoff[i] = (int64_t)i * ne;
    for (int i = 0; i < ni; i++) c->nunk[i] = b->n_unknowns[0];
  } else {
    if (!b->elem_offset) return set_err(c, KG_INVALID_ARGUMENT, "elem_offset required for per-instance topologies");
    for (int i = 0; i <= ni; i++) c->eoff[i] = b->elem_offset[i];
    c->el.assign(b->elements, b->elements + b->elem_offset[ni]);
    for (int i = 0; i < ni; i++) c->nunk[i] = b->n_unknowns[i];
  }
  int64_t nrec = c->shared ? (int64_t)ni * b->n_elements : c->eoff[ni];
  c->params.assign(b->params, b->params + nrec * KG_NPARAM);
  c->ebase.resize(nrec);
  for (int i = 0; i < ni; i++) {
    int64_t e0 = c->eoff[i], e1 = c->eoff[i + 1];
    for (int64_t r = e0; r < e1; r++) {
      const kg_element& e = c->el[c->shared ? r - e0 : r];
      double v = c->params[r * KG_NPARAM];
      c->ebase[r] = (e.type == KG_R && !e.g2) ? 1.0 / v : v;  // stampable.rs:232 g = 1.0 / value
    }
  }
  if (c->shared) {
    // element records are per instance in params/ebase; the element table is shared
    std::vector<kg_element> rep(c->el);
    c->el.clear();
    for (int i = 0; i < ni; i++) c->el.insert(c->el.end(), rep.begin(), rep.end());
    c->group_of.assign(ni, 0);
    c->group_first.assign(1, 0);
  } else {
    // group identical topologies
    std::map<std::vector<int32_t>, int> sig;
    c->group_of.assign(ni, 0);
    c->group_first.clear();
    for (int i = 0; i < ni; i++) {
      std::vector<int32_t> s{c->nunk[i]};
      for (int64_t r = c->eoff[i]; r < c->eoff[i + 1]; r++) {
        const kg_element& e = c->el[r];
        s.insert(s.end(), {e.type, e.plus, e.minus, e.branch, e.g2, e.ctrl_plus, e.ctrl_minus, e.ctrl_branch});
      }
      auto it = sig.find(s);
      if (it == sig.end()) {
        int g = (int)c->group_first.size();
        sig[s] = g;
        c->group_first.push_back(i);
        c->group_of[i] = g;
      } else {
        c->group_of[i] = it->second;
      }
    }
  }
  for (int i = 0; i < ni; i++)
    for (int64_t r = c->eoff[i]; r < c->eoff[i + 1]; r++) {
      const kg_element& e = c->el[r];
      int n = c->nunk[i];
      if (e.type < KG_R || e.type > KG_H || e.plus >= n || e.minus >= n || e.branch >= n)
        return set_err(c, KG_INVALID_ELEMENT_FORMAT, "element table out of range");
      if ((e.type == KG_V || e.type == KG_L || e.type == KG_I || e.type == KG_E || e.type == KG_H) && e.branch < 0)
        return set_err(c, KG_INVALID_ELEMENT_FORMAT, "V, L, I, E and H elements need a branch unknown");
      if ((e.type == KG_E || e.type == KG_G) && (e.ctrl_plus >= n || e.ctrl_minus >= n))
        return set_err(c, KG_INVALID_ELEMENT_FORMAT, "controlling node out of range");
      if ((e.type == KG_F || e.type == KG_H) && (e.ctrl_branch < 0 || e.ctrl_branch >= n))
        return set_err(c, KG_INVALID_ELEMENT_FORMAT, "F and H elements need a controlling branch unknown");
    }
  HIPCHK(c, hipMalloc(&c->d_params, sizeof(double) * std::max<size_t>(c->params.size(), 1)));
  HIPCHK(c, hipMalloc(&c->d_ebase, sizeof(double) * std::max<size_t>(c->ebase.size(), 1)));
  HIPCHK(c, hipMalloc(&c->d_eoff, sizeof(int64_t) * c->eoff.size()));
  HIPCHK(c, hipMalloc(&c->d_nunk, sizeof(int32_t) * c->nunk.size()));
  HIPCHK(c, hipMemcpy(c->d_params, c->params.data(), sizeof(double) * c->params.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_ebase, c->ebase.data(), sizeof(double) * c->ebase.size(), hipMemcpyHostToDevice));
  if (c->shared) {
    const int64_t ne = b->n_elements;
    std::vector<double> t((size_t)((ni + 63) / 64) * 64 * ne, 0.0);
    for (int64_t i = 0; i < ni; i++)
      for (int64_t e = 0; e < ne; e++) t[(i / 64 * ne + e) * 64 + i % 64] = c->ebase[i * ne + e];
    HIPCHK(c, hipMalloc(&c->d_ebase_t, sizeof(double) * std::max<size_t>(t.size(), 1)));
    HIPCHK(c, hipMemcpy(c->d_ebase_t, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpy(c->d_eoff, c->eoff.data(), sizeof(int64_t) * c->eoff.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_nunk, c->nunk.data(), sizeof(int32_t) * c->nunk.size(), hipMemcpyHostToDevice));
  return KG_OK;
}

int32_t kg_op(kg_ctx* c, const kg_solver_config* cfg, double* x, int32_t* status, int32_t* iters, int32_t mem) {
  if (!c || !cfg || c->n_inst == 0) return set_err(c, KG_INVALID_ARGUMENT, "no batch loaded");
  c->err.clear();
  HIPCHK(c, hipSetDevice(c->device));
  double* d_out;
  int64_t* d_off;
  int32_t *d_st, *d_it;
  int32_t r = run_dc(c, cfg, -1, 0.0, 0.0, 1, &d_out, &d_off, &d_st, &d_it);
  if (r) return r;
  int64_t tot = 0;
  for (int i = 0; i < c->n_inst; i++) tot += c->nunk[i];
  return finish_status(c, d_st, status, mem,
                       {{x, d_out, sizeof(double) * tot, true}, {iters, d_it, sizeof(int32_t) * c->n_inst, false}});
}

int32_t kg_dc_sweep(kg_ctx* c, const kg_solver_config* cfg, int32_t element, double start, double stop, double step,
                    double* x, int32_t max_points, int32_t* n_points, int32_t* status, int32_t* iters, int32_t mem) {
  if (!c || !cfg || c->n_inst == 0) return set_err(c, KG_INVALID_ARGUMENT, "no batch loaded");
  c->err.clear();
  HIPCHK(c, hipSetDevice(c->device));
  // dc.rs:23-38: the swept element must exist and be a V or I source
  for (int i = 0; i < c->n_inst; i++) {
    int64_t ne = c->eoff[i + 1] - c->eoff[i];
    if (element < 0 || element >= ne) return set_err(c, KG_ELEMENT_NOT_FOUND, "swept element not found");
    int ty = c->el[c->eoff[i] + element].type;
    if (ty != KG_V && ty != KG_I)
      return set_err(c, KG_INVALID_ELEMENT_FORMAT, "DC sweep element must be a voltage or current source");
  }
  // dc.rs:130-131
  double nf = std::fabs((stop - start) / step);
  if (!(nf < 1e9)) return set_err(c, KG_INVALID_ARGUMENT, "invalid sweep range");
  int npts = (int)(int64_t)nf + 1;
  if (n_points) *n_points = npts;
  if (npts > max_points) return set_err(c, KG_INVALID_ARGUMENT, "output buffer too small for the sweep");
  double* d_out;
  int64_t* d_off;
  int32_t *d_st, *d_it;
  int32_t r = run_dc(c, cfg, element, start, step, npts, &d_out, &d_off, &d_st, &d_it);
  if (r) return r;
  int64_t tot = 0;
  for (int i = 0; i < c->n_inst; i++) tot += (int64_t)c->nunk[i] * npts;
  return finish_status(c, d_st, status, mem,
                       {{x, d_out, sizeof(double) * tot, true},
                        {iters, d_it, sizeof(int32_t) * c->n_inst * (size_t)npts, false}});
}

int32_t kg_ac_sweep(kg_ctx* c, const kg_solver_config* cfg, const double* freqs, int32_t nf, int32_t precision,
                    double* x, int32_t* status, double* err_est, int32_t mem) {
  if (!c || !cfg || c->n_inst == 0 || !freqs || nf <= 0) return set_err(c, KG_INVALID_ARGUMENT, "bad arguments");
  c->err.clear();
  HIPCHK(c, hipSetDevice(c->device));
  for (int k = 0; k < nf; k++)
    if (!(freqs[k] > 0.0)) return set_err(c, KG_INVALID_ARGUMENT, "frequencies must be positive (ac.rs:40-45)");
  Plan P0, P;
  double fmin = freqs[0], fmax = freqs[0];
  for (int k = 0; k < nf; k++) fmin = std::min(fmin, freqs[k]), fmax = std::max(fmax, freqs[k]);
  double s_ac = 2.0 * 3.14159265358979323846 * std::sqrt(fmin * fmax);
  int32_t r = prepare(c, AK_AC, ORD_NESTED, s_ac, P0);
  if (r) return r;
  const std::string tkey = tune_key(c, "ac", precision, P0, (int64_t)c->n_inst * nf);
  // A call whose work mapping is already known runs as one captured graph:
  // operating point (fill, Newton kernel), fill, sweep kernel.
  bool graph = c->graphs && !c->prof && !c->noalloc && (cfg->threads_per_system != 0 || c->tune.count(tkey));
  std::vector<Deferred> steps;
  // ac.rs:25: the operating point comes first
  kg_solver_config ocfg = *cfg;
  ocfg.threads_per_system = 0;
  double* d_xop;
  int64_t* d_xoff;
  int32_t *d_st, *d_it;
  r = run_dc(c, &ocfg, -1, 0.0, 0.0, 1, &d_xop, &d_xoff, &d_st, &d_it, graph ? &steps : nullptr);
  if (r) return r;
  if (steps.empty()) graph = false;  // the operating point ran eagerly (structurally singular groups)
  bool mixed = precision == KG_MIXED;
  int refine = cfg->refine_iters > 0 ? cfg->refine_iters : 2;
  auto off = out_offsets(c, 2 * (int64_t)nf);
  int64_t* d_off;
  if ((r = upload_offsets(c, "ac_off", off, &d_off))) return r;
  hipError_t e;
  double* d_f = nullptr;
  if ((r = upload(c, "freqs", freqs, sizeof(double) * nf, (void**)&d_f))) return r;
  double* d_out = (mem == KG_MEM_DEVICE && x) ? x : (double*)ws_get(c, "ac_out", sizeof(double) * off.back(), &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  double* d_err = (double*)ws_get(c, "err", sizeof(double) * c->n_inst, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  // the accuracy guard's queue: a count, then up to one entry per system
  int64_t fixcap = (int64_t)c->n_inst * nf;
  if (c->noalloc) {  // the queue's region of the scratch array bounds it
    auto it = c->given.find("fixq");
    if (it != c->given.end()) fixcap = std::min<int64_t>(fixcap, (int64_t)(it->second.second / 8) - 1);
  }
  auto* d_fix = (unsigned long long*)ws_get(c, "fixq", sizeof(unsigned long long) * (1 + (size_t)fixcap), &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  // One AC launch for a mapping candidate: TPS threads per system, plus 1000
  // for the persistent variant (shared topology only) that stages the plan
  // and each instance's element records in LDS.
  // 2001: the lane-per-system kernel specialised to this topology at run
  // time (shared topology, small systems); 2002: its mixed precision form
  // with the solution split into a complex64 first solution in LDS and a
  // complex64 sum of corrections in registers. 2011 / 2012: the same on the
  // balanced order (independent elimination chains, more live registers).
  struct AcLaunch {
    Plan P;
    Lay lay;
    Launch L;
    bool persistent = false, lane = false;
    int tps = 0, plan_ints = 0;
    int64_t chunk = 0;   // persistent: systems per block
    int64_t gchunk = 0;  // systems larger than LDS: systems per launch
    JitFn fn;
  };
  const int64_t nsys = (int64_t)c->n_inst * nf;
  auto setup = [&](int cand, AcLaunch& A) -> int32_t {
    A.lane = cand == 2001 || cand == 2002 || cand == 2011 || cand == 2012;
    A.persistent = cand >= 1000 && !A.lane;
    A.tps = cand % 1000;
    if (A.lane) {
      const bool split = cand % 10 == 2;
      const Ordering ord = cand >= 2010 ? ORD_BALANCED : ORD_MINDEG;
      if (!c->shared) return set_err(c, KG_UNEXPECTED, "lane kernel needs a shared topology");
      int32_t rr = prepare(c, AK_AC, ord, s_ac, A.P);
      if (rr) return rr;
      if (split && !mixed) return set_err(c, KG_UNEXPECTED, "2002 / 2012 are mixed precision lane kernels");
      int pidx = c->pat_index[std::make_tuple(0, (int)AK_AC, (int)ord, s_ac)];
      auto key = std::make_tuple(pidx, (int)mixed, (int)split);
      auto it = c->lane_fn.find(key);
      if (it == c->lane_fn.end()) {
        if (c->noalloc) return set_err(c, KG_INVALID_ARGUMENT, "allocation-free call: kernel not compiled yet");
        const PatEntry& pe = c->pats[pidx];
        std::string src, why, log;
        if (!jit_ac_source(pe.lay, pe.host.data(), mixed, split, src, why))
          return set_err(c, KG_UNEXPECTED, "lane kernel: " + why);
        if (jit_load(c->device, src, jit_ac_name(mixed, split), A.fn, log) != hipSuccess)
          return set_err(c, KG_UNEXPECTED, "lane kernel compilation failed: " + log);
        c->lane_fn[key] = A.fn;
      } else {
        A.fn = it->second;
      }
      A.L.grid = dim3((unsigned)((nsys + 63) / 64));
      A.L.block = dim3(64);
      A.L.lds = split ? 64 * sizeof(float) * 2 * (size_t)c->nunk[0] : 0;  // x0 per lane (complex64)
      return KG_OK;
    }
    if (!valid_tps(A.tps)) return set_err(c, KG_INVALID_ARGUMENT, "threads_per_system must be 1, 4, 16, 64 or 256");
    int32_t rr = prepare(c, AK_AC, order_for(A.tps), s_ac, A.P);
    if (rr) return rr;
    A.lay = Lay{A.P.nnz_max, A.P.n_max, A.P.ne_max, 0, 0, A.P.any_diode ? 1 : 0, mixed ? 8 : 16, 16, mixed ? 1 : 0, 0};
    if (!A.persistent) {
      const int32_t rc = configure(c, A.tps, A.lay, nsys, A.L);
      if (rc == KG_OK || A.tps < 64) return rc;
      c->err.clear();  // too large for LDS: the values go to a global workspace
      return configure_global(c, A.tps, A.lay, nsys, A.L, "ac_gws", &A.gchunk);
    }
    if (!c->shared) return set_err(c, KG_UNEXPECTED, "persistent AC mapping needs a shared topology");
    const PatEntry& pe = c->pats[c->pat_index[std::make_tuple(0, (int)AK_AC, (int)order_for(A.tps), s_ac)]];
    A.plan_ints = pe.lay.total;
    size_t fixed = lay_al(4 * (size_t)A.plan_ints) + lay_al(8 * (size_t)A.P.ne_max * KG_NPARAM) +
                   lay_al(8 * (size_t)A.P.ne_max);
    size_t sb = sys_bytes(A.lay);
    int spb = std::max(1, 64 / A.tps);
    size_t lds = fixed + spb * sb + 8 * (size_t)spb * A.tps;
    if (lds > LDS_MAX) return set_err(c, KG_UNEXPECTED, "persistent AC mapping does not fit in LDS");
    int per_cu = std::max<int>(1, (int)std::min<size_t>(8, LDS_MAX / lds));
    int64_t items = (int64_t)c->n_inst * ((nf + spb - 1) / spb);
    int64_t grid = std::min<int64_t>(items, 256 * (int64_t)per_cu);
    A.chunk = (items + grid - 1) / grid;
    grid = (items + A.chunk - 1) / A.chunk;
    A.L.tps = A.tps;
    A.L.spb = spb;
    A.L.lds = lds;
    A.L.block = dim3(A.tps * spb);
    A.L.grid = dim3((unsigned)grid);
    return KG_OK;
  };
  auto run = [&](const AcLaunch& A) -> hipError_t {
    if (A.lane) {
      const double* prm = c->d_params;
      const double* eb = mixed ? c->d_ebase_t : c->d_ebase;  // mixed: blocks of 64 instances
      const double* fr = d_f;
      int nfa = nf;
      int64_t ns = nsys;
      double* outp = d_out;
      int32_t* stp = d_st;
      double* errp = d_err;
      int rf = refine;
      double stop = cfg->refine_iters > 0 ? 0.0 : KG_REFINE_STOP;  // explicit counts run every pass
      unsigned long long* fq = d_fix;
      int64_t fc = fixcap;
      void* args[] = {&prm, &eb, &fr, &nfa, &ns, &outp, &stp, &errp, &rf, &stop, &fq, &fc};
      return hipModuleLaunchKernel(A.fn.fn, A.L.grid.x, 1, 1, 64, 1, 1, (unsigned)A.L.lds, c->stream, args, nullptr);
    }
    BatchDev B = batch_dev(c, A.P);
    if (A.persistent)
      return run_ac_persistent(A.tps, mixed ? 1 : 0, A.L.grid, A.L.block, A.L.lds, c->stream, B, 0, d_f, nf, d_xop,
                               d_xoff, d_out, d_off, d_st, d_err, refine, A.lay, A.plan_ints, A.chunk, d_fix, fixcap);
    return chunked(A.lay, nsys, A.gchunk, A.L, [&](Lay ly, dim3 g) {
      return run_ac(A.tps, mixed ? 1 : 0, g, A.L.block, A.L.lds, c->stream, B, 0, d_f, nf, d_xop, d_xoff, d_out, d_off,
                    d_st, d_err, refine, ly, d_fix, fixcap);
    });
  };
  // the guard's fixup over whatever the sweep kernel queued (persistent grid):
  // systems of up to nmax unknowns in LDS, larger ones (up to KG_FIX_GMAX) in
  // a second launch with a global workspace region per block (fix_gws, set
  // up with the final mapping)
  char* fix_gws = nullptr;
  int64_t fix_gstride = 0;
  int fix_gmax = 0;
  auto fix_nmax = [&](const AcLaunch& A) {
    int nmax = std::min(A.P.n_max, KG_FIX_NMAX);
    while (nmax > 1 && ac_fix_lds(nmax, A.P.ne_max) > LDS_MAX) nmax--;
    return nmax;
  };
  auto fixup = [&](const AcLaunch& A) -> hipError_t {
    const int nmax = fix_nmax(A);
    const bool big = fix_gws && A.P.n_max > nmax;
    const BatchDev B = batch_dev(c, A.P);
    hipError_t he = run_ac_fix(256, c->stream, B, 0, d_f, nf, d_xop, d_xoff, d_out, d_off, d_st, d_err,
                               mixed ? 1 : 0, d_fix, fixcap, nmax, A.P.ne_max, 0, big ? 1 : 0, nullptr, 0);
    if (he != hipSuccess || !big) return he;
    return run_ac_fix(KG_FIX_GBLOCKS, c->stream, B, 0, d_f, nf, d_xop, d_xoff, d_out, d_off, d_st, d_err,
                      mixed ? 1 : 0, d_fix, fixcap, fix_gmax, A.P.ne_max, nmax, 0, fix_gws, fix_gstride);
  };
  int cand = cfg->threads_per_system;
  if (cand == 0 && c->noalloc && !c->tune.count(tkey))
    return set_err(c, KG_INVALID_ARGUMENT,
                   "allocation-free call: the work mapping is not known yet (give threads_per_system, or call "
                   "kg_ac_sweep once on this workload first)");
  if (cand == 0) {
    std::vector<int> cands{1, 4, 16, 64, 256};
    if (c->shared) cands.insert(cands.end(), {1001, 1004, 1016, 1064, 2001, 2011});
    if (c->shared && mixed) cands.insert(cands.end(), {2002, 2012});
    r = autotune(c, tkey, cands, [&](int t, float* ms) -> int32_t {
      AcLaunch A;
      int32_t rr = setup(t, A);
      if (rr) return rr;
      if (!A.lane && A.tps * A.L.spb < 64 && A.tps < 64)
        return set_err(c, KG_UNEXPECTED, "mapping leaves most of a wave idle");
      HIPCHK(c, run_fill_f64((double*)d_fix, 0.0, 1, c->stream));
      return time_launch(c, [&] { return run(A); }, ms);
    }, &cand);
    if (r) return r;
  }
  AcLaunch A;
  if ((r = setup(cand, A))) return r;
  P = A.P;
  if (A.P.n_max > fix_nmax(A) && !c->noalloc) {  // queued systems above the LDS fixup's size
    fix_gmax = std::min(A.P.n_max, KG_FIX_GMAX);
    fix_gstride = (int64_t)lay_al(ac_fix_lds(fix_gmax, A.P.ne_max));
    hipError_t eg;
    fix_gws = (char*)ws_get(c, "fix_gws", (size_t)fix_gstride * KG_FIX_GBLOCKS, &eg);
    if (eg != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (fixup regions)");
  }
  int ni = c->n_inst;
  // error figures and the guard's queue count, zeroed in one launch
  steps.push_back({nullptr, [&] { return run_zero2_f64(d_err, ni, (double*)d_fix, 1, c->stream); }});
  steps.push_back({A.lane ? jit_ac_name(mixed, cand % 10 == 2)
                   : A.persistent ? (mixed ? "k_ac_p<complex64,complex128>" : "k_ac_p<complex128,complex128>")
                                  : (mixed ? "k_ac<complex64,complex128>" : "k_ac<complex128,complex128>"),
                   [&] { return run(A); }});
  steps.push_back({"k_ac_fix", [&] { return fixup(A); }});
  if (graph) {
    std::ostringstream k;
    k << "ac|" << cand << "|" << refine << "|" << (cfg->refine_iters > 0 ? "fixed" : "adaptive") << "|" << precision
      << "|" << nf << "|" << ocfg.reltol << "|" << ocfg.iabstol
      << "|" << ocfg.vabstol << "|" << ocfg.maxiter << "|" << ocfg.compat << "|" << (void*)d_out << "|" << (void*)d_f
      << "|" << (void*)d_xop << "|" << (void*)d_st << "|" << (void*)d_err << "|" << (void*)c->d_pats << "|"
      << (void*)d_off << "|" << (void*)d_xoff << "|" << (void*)A.fn.fn << "|" << (void*)d_fix;
    auto it = c->gexec.find(k.str());
    if (it == c->gexec.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t ge = nullptr;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e2 = hipSuccess;
      for (auto& d : steps)
        if ((e2 = d.launch()) != hipSuccess) break;
      hipError_t e3 = hipStreamEndCapture(c->stream, &g);
      if (e2 != hipSuccess || e3 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return set_err(c, KG_HIP_ERROR, std::string("graph capture: ") + hipGetErrorString(e2 ? e2 : e3));
      }
      HIPCHK(c, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      it = c->gexec.emplace(k.str(), std::make_pair(g, ge)).first;
    }
    HIPCHK(c, hipGraphLaunch(it->second.second, c->stream));
  } else {
    for (auto& d : steps) {
      if (d.name) {
        if ((r = timed(c, d.name, d.launch))) return r;
      } else {
        HIPCHK(c, d.launch());
      }
    }
  }
  if ((r = mark_group_failures(c, P, d_st))) return r;
  return finish_status(c, d_st, status, mem,
                       {{x, d_out, sizeof(double) * off.back(), true}, {err_est, d_err, sizeof(double) * c->n_inst, false}});
}

// Allocation-free AC sweep: kg_ac_sweep with every intermediate in the
// device's scratch array and the outputs in the caller's device buffers.
int32_t kg_ac_sweep_noalloc(kg_ctx* c, const kg_solver_config* cfg, const double* freqs, int32_t nf,
                            int32_t precision, double* x, int32_t* status, double* err_est) {
  if (!c || !cfg || c->n_inst == 0 || !freqs || nf <= 0 || !x || !status)
    return set_err(c, KG_INVALID_ARGUMENT, "bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  char* base = c->device < KG_MAX_DEVICES ? (char*)g_scratch_ptr[c->device] : nullptr;
  if (!base)
    return set_err(c, KG_HIP_ERROR,
                   std::string("scratch array unavailable on this device: ") +
                       hipGetErrorString(c->device < KG_MAX_DEVICES ? g_scratch_err[c->device] : hipErrorInvalidDevice));
  if (!on_device(x) || !on_device(status) || (err_est && !on_device(err_est)))
    return set_err(c, KG_INVALID_ARGUMENT, "x, status and err_est must be device memory of the context's GPU");
  const int64_t ni = c->n_inst;
  int64_t ntot = 0;
  for (int32_t v : c->nunk) ntot += v;
  size_t used = 0;
  std::map<std::string, std::pair<void*, size_t>> reg;
  auto carve = [&](const char* name, size_t bytes) {
    used = (used + 255) & ~(size_t)255;
    reg[name] = {base + used, bytes};
    used += bytes;
  };
  carve("newton_out", 8 * (size_t)std::max<int64_t>(ntot, 1));
  carve("newton_off", 8 * (size_t)(ni + 1));
  carve("newton_iters", 4 * (size_t)ni);
  carve("ac_off", 8 * (size_t)(ni + 1));
  carve("freqs", 8 * (size_t)nf);
  if (!err_est) carve("err", 8 * (size_t)ni);
  carve("fixq", 8 * (size_t)(1 + std::min<int64_t>(ni * nf, KG_NOALLOC_FIXQ)));
  if (used > KG_NOALLOC_SCRATCH_BYTES)
    return set_err(c, KG_INVALID_ARGUMENT,
                   "allocation-free call needs " + std::to_string(used) + " bytes of scratch; the maximum is " +
                       std::to_string(KG_NOALLOC_SCRATCH_BYTES));
  reg["status"] = {status, 4 * (size_t)ni};
  if (err_est) reg["err"] = {err_est, 8 * (size_t)ni};
  struct Reset {
    kg_ctx* c;
    ~Reset() {
      c->noalloc = false;
      c->given.clear();
    }
  } reset{c};
  c->given = std::move(reg);
  c->noalloc = true;
  return kg_ac_sweep(c, cfg, freqs, nf, precision, x, status, err_est, KG_MEM_DEVICE);
}

int32_t kg_transient(kg_ctx* c, const kg_solver_config* cfg, double h, int32_t nsteps, int32_t method,
                     int32_t precision, double* x, int32_t* status, int32_t* iters, double* err_est, int32_t mem) {
  if (!c || !cfg || c->n_inst == 0 || !(h > 0) || nsteps < 0 || method < KG_BE_COMPAT || method > KG_TRAP)
    return set_err(c, KG_INVALID_ARGUMENT, "bad arguments");
  c->err.clear();
  HIPCHK(c, hipSetDevice(c->device));
  kg_solver_config ocfg = *cfg;
  ocfg.threads_per_system = 0;
  ocfg.compat = method == KG_BE_COMPAT ? 1 : 0;  // transient.rs:21 op::solve; clean mode starts from zero
  if (method != KG_BE_COMPAT) ocfg.compat = cfg->compat && method == KG_BE_COMPAT;
  double* d_xop;
  int64_t* d_xoff;
  int32_t *d_st, *d_it;
  int32_t r = run_dc(c, &ocfg, -1, 0.0, 0.0, 1, &d_xop, &d_xoff, &d_st, &d_it);
  if (r) return r;
  Plan P0, P;
  double s_tr = (method == KG_TRAP ? 2.0 : 1.0) / h;
  if ((r = prepare(c, AK_TRAN, ORD_NESTED, s_tr, P0))) return r;
  bool mixed = precision == KG_MIXED;
  auto off = out_offsets(c, (int64_t)nsteps + 1);
  int64_t* d_off;
  if ((r = upload_offsets(c, "tran_off", off, &d_off))) return r;
  hipError_t e;
  double* d_out = (mem == KG_MEM_DEVICE && x) ? x : (double*)ws_get(c, "tran_out", sizeof(double) * off.back(), &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  int32_t* d_iters = (int32_t*)ws_get(c, "tran_iters", sizeof(int32_t) * (size_t)c->n_inst * (nsteps + 1), &e);
  double* d_err = (double*)ws_get(c, "err", sizeof(double) * c->n_inst, &e);
  int32_t* d_st0 = (int32_t*)ws_get(c, "tran_status0", sizeof(int32_t) * c->n_inst, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  // the kernel overwrites the status array: keep the operating-point statuses
  HIPCHK(c, hipMemcpyAsync(d_st0, d_st, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToDevice, c->stream));
  TranArgs ta{h, method == KG_TRAP ? 2.0 : 1.0, nsteps, method, cfg->refine_iters > 0 ? cfg->refine_iters : 2};
  NewtonArgs na = newton_args(cfg);
  na.compat = method == KG_BE_COMPAT ? 1 : 0;
  // cand: threads per system, plus 1000 for the plan staged in LDS (64, 256).
  // chunk: systems per launch (systems larger than LDS, whose values go to a
  // global workspace with 64 or 256 threads each).
  auto setup = [&](int cand, Plan& PP, Lay& lay, Launch& L, int& pi, int64_t& chunk) -> int32_t {
    int tps = cand % 1000;
    int32_t rr = prepare(c, AK_TRAN, order_for(tps), s_tr, PP);
    if (rr) return rr;
    lay = Lay{PP.nnz_max, PP.n_max, PP.ne_max, 1, 1, 1, mixed ? 4 : 8, 8, mixed ? 1 : 0, 0};
    pi = cand >= 1000 ? PP.total_max : 0;
    if (pi && tps != 64 && tps != 256) return set_err(c, KG_INVALID_ARGUMENT, "staged plans need 64 or 256 threads");
    chunk = c->n_inst;
    rr = configure(c, tps, lay, c->n_inst, L, pi ? lay_al(4 * (size_t)pi) : 0);
    if (rr == KG_OK || pi || tps < 64) return rr;
    c->err.clear();
    return configure_global(c, tps, lay, c->n_inst, L, "tran_gws", &chunk);
  };
  if (!c->shared && cfg->threads_per_system == 0 && c->n_inst >= 64) {
    r = run_tran_buckets(c, na, ta, mixed, s_tr, precision, d_xop, d_xoff, d_out, d_off, d_st, d_st0, d_iters, d_err);
    if (r) return r;
    P = P0;
  } else {
  int cand = cfg->threads_per_system;
  // k_tran_lin (3000 + threads per block): FP64, linear topologies with a compact plan
  size_t lin = 0;
  auto lin_setup = [&](Plan& PP) -> int32_t {
    int32_t rr = prepare(c, AK_TRAN, ORD_NESTED, s_tr, PP);
    if (rr) return rr;
    lin = mixed ? 0 : lin_need_all(c, s_tr);
    if (lin == 0 || lin > LDS_MAX)
      return set_err(c, KG_UNEXPECTED, "the LDS-resident kernel needs FP64 and a linear topology that fits LDS");
    return KG_OK;
  };
  auto lin_launch = [&](int cand, const Plan& PP, hipStream_t st) -> hipError_t {
    const int rows = lin_rows(cand - 3000, PP.n_max);
    if (rows == 0) return hipErrorInvalidConfiguration;
    int nsrc_max = 0;
    int32_t rr = KG_OK;
    double* srcv = lin_srcv(c, s_tr, nsteps, &nsrc_max, &rr);
    if (rr) return hipErrorOutOfMemory;
    return run_tran_lin(cand - 3000, rows, dim3((unsigned)c->n_inst), lin, st, batch_dev(c, PP), 0, ta, d_xop,
                        d_xoff, d_out, d_off, d_st, d_iters, d_err, srcv, nsrc_max);
  };
  // 2001: one lane per instance, code specialised to the shared topology at
  // run time (FP64; diodes included)
  JitFn tran_fn;
  auto lane_setup = [&](Plan& PP) -> int32_t {
    if (!c->shared || mixed) return set_err(c, KG_UNEXPECTED, "the lane kernel needs FP64 and a shared topology");
    if (method != KG_BE_COMPAT && any_junction_charge(c))
      return set_err(c, KG_UNEXPECTED, "lane kernel: junction charge is not generated");
    int32_t rr = prepare(c, AK_TRAN, ORD_MINDEG, s_tr, PP);
    if (rr) return rr;
    if (PP.group_status[0] != KG_OK) return set_err(c, KG_UNEXPECTED, "lane kernel: no valid plan");
    const int pidx = c->pat_index[std::make_tuple(0, (int)AK_TRAN, (int)ORD_MINDEG, s_tr)];
    const auto key = std::make_tuple(pidx, 300, 0);
    auto it = c->lane_fn.find(key);
    if (it != c->lane_fn.end()) {
      tran_fn = it->second;
      return tran_fn.fn ? KG_OK : set_err(c, KG_UNEXPECTED, "lane kernel: topology not specialised");
    }
    const PatEntry& pe = c->pats[pidx];
    std::string src, why, log;
    if (!jit_tran_source(pe.lay, pe.host.data(), src, why)) {
      c->lane_fn[key] = JitFn{};
      return set_err(c, KG_UNEXPECTED, "lane kernel: " + why);
    }
    if (jit_load(c->device, src, jit_tran_name(), tran_fn, log) != hipSuccess)
      return set_err(c, KG_UNEXPECTED, "lane kernel compilation failed: " + log);
    c->lane_fn[key] = tran_fn;
    return KG_OK;
  };
  auto lane_launch = [&]() -> hipError_t {
    const double* prm = c->d_params;
    const double* eb = c->d_ebase;
    int a_ni = c->n_inst, a_ns = nsteps, a_m = method, a_mi = na.maxiter;
    double a_h = h, a_al = ta.alpha, a_rt = na.reltol, a_it = na.iabstol, a_vt = na.vabstol;
    void* args[] = {&prm, &eb, &d_xop, &d_xoff, &d_out, &d_off, &d_st, &d_iters, &d_err,
                    &a_ni, &a_h, &a_al, &a_ns, &a_m, &a_rt, &a_it, &a_vt, &a_mi};
    return hipModuleLaunchKernel(tran_fn.fn, (unsigned)((c->n_inst + 63) / 64), 1, 1, 64, 1, 1, 0, c->stream, args,
                                 nullptr);
  };
  if (cand == 0) {
    std::vector<int> cands = {1, 4, 16, 64, 256, 1064, 1256};
    cands.insert(cands.end(), LIN_CANDS.begin(), LIN_CANDS.end());
    if (c->shared && !mixed) cands.push_back(2001);
    r = autotune(c, tune_key(c, "tran", precision, P0, c->n_inst), cands, [&](int t, float* ms) -> int32_t {
      Plan PP;
      if (t == 2001) {
        int32_t rr = lane_setup(PP);
        if (rr) return rr;
        HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToDevice, c->stream));
        return time_launch(c, [&] { return lane_launch(); }, ms);
      }
      if (is_lin_cand(t)) {
        int32_t rr = lin_setup(PP);
        if (rr) return rr;
        if (lin_rows(t - 3000, PP.n_max) == 0) return set_err(c, KG_UNEXPECTED, "more than 4 rows per thread");
        HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToDevice, c->stream));
        return time_launch(c, [&] { return lin_launch(t, PP, c->stream); }, ms);
      }
      Lay lay;
      Launch L;
      int pi = 0;
      int64_t chunk = 0;
      int32_t rr = setup(t, PP, lay, L, pi, chunk);
      if (rr) return rr;
      if ((t % 1000) * L.spb < 64 && (t % 1000) < 64) return set_err(c, KG_UNEXPECTED, "sub-wave mapping");
      BatchDev B = batch_dev(c, PP);
      HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToDevice, c->stream));
      return time_launch(c, [&] {
        return chunked(lay, c->n_inst, chunk, L, [&](Lay ly, dim3 g) {
          return run_tran(t % 1000, mixed ? 1 : 0, g, L.block, L.lds, c->stream, B, 0, na, ta, d_xop, d_xoff, d_out,
                          d_off, d_st, d_iters, d_err, ly, pi);
        });
      }, ms);
    }, &cand);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * c->n_inst, hipMemcpyDeviceToDevice, c->stream));
  }
  if (cand == 2001) {
    if ((r = lane_setup(P))) return r == KG_UNEXPECTED ? set_err(c, KG_INVALID_ARGUMENT, c->err) : r;
    r = timed(c, jit_tran_name(), [&] { return lane_launch(); });
    if (r) return r;
  } else if (is_lin_cand(cand)) {
    if ((r = lin_setup(P))) return r == KG_UNEXPECTED ? set_err(c, KG_INVALID_ARGUMENT, c->err) : r;
    if (lin_rows(cand - 3000, P.n_max) == 0)
      return set_err(c, KG_INVALID_ARGUMENT, "system too large for this block size (more than 4 rows per thread)");
    r = timed(c, "k_tran_lin", [&] { return lin_launch(cand, P, c->stream); });
    if (r) return r;
  } else {
  if (!valid_tps(cand % 1000) || (cand >= 1000 && cand != 1064 && cand != 1256))
    return set_err(c, KG_INVALID_ARGUMENT,
                   "threads_per_system must be 1, 4, 16, 64, 256, 1064, 1256, 2001, 3064, 3128, 3256, 3512 or 4024");
  Lay lay;
  Launch L;
  int pi = 0;
  int64_t chunk = 0;
  if ((r = setup(cand, P, lay, L, pi, chunk))) return r;
  BatchDev B = batch_dev(c, P);
  r = timed(c, mixed ? "k_tran<float,double>" : "k_tran<double,double>", [&] {
    return chunked(lay, c->n_inst, chunk, L, [&](Lay ly, dim3 g) {
      return run_tran(cand % 1000, mixed ? 1 : 0, g, L.block, L.lds, c->stream, B, 0, na, ta, d_xop, d_xoff, d_out,
                      d_off, d_st, d_iters, d_err, ly, pi);
    });
  });
  if (r) return r;
  }
  }
  if ((r = mark_group_failures(c, P, d_st))) return r;
  return finish_status(c, d_st, status, mem,
                       {{x, d_out, sizeof(double) * off.back(), true},
                        {iters, d_iters, sizeof(int32_t) * (size_t)c->n_inst * (nsteps + 1), false},
                        {err_est, d_err, sizeof(double) * c->n_inst, false}});
}

int32_t kg_last_autotune(const kg_ctx* c, int32_t* tps, float* ms, int32_t cap) {
  if (!c) return -KG_INVALID_ARGUMENT;
  int32_t n = (int32_t)c->last_trials.size();
  for (int32_t k = 0; k < std::min(n, cap); k++) {
    if (tps) tps[k] = c->last_trials[k].first;
    if (ms) ms[k] = c->last_trials[k].second;
  }
  return n;
}

int32_t kg_use_graphs(kg_ctx* c, int32_t enable) {
  if (!c) return KG_INVALID_ARGUMENT;
  c->graphs = enable != 0;
  return KG_OK;
}

int32_t kg_profile(kg_ctx* c, int32_t enable, double* total_ms, int64_t* launches, const char** kernel) {
  if (!c) return KG_INVALID_ARGUMENT;
  // the dominant kernel: largest accumulated time
  double ms = 0.0;
  int64_t n = 0;
  c->prof_name.clear();
  for (auto& kv : c->prof_k)
    if (kv.second.first > ms) {
      ms = kv.second.first;
      n = kv.second.second;
      c->prof_name = kv.first;
    }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = n;
  if (kernel) *kernel = c->prof_name.c_str();
  c->prof = enable != 0;
  c->prof_k.clear();
  return KG_OK;
}

int32_t kg_profile_summary(const kg_ctx* c, char* buf, int32_t cap) {
  if (!c) return -KG_INVALID_ARGUMENT;
  std::string out;
  char line[64];
  for (auto& kv : c->prof_k) {
    std::snprintf(line, sizeof(line), " %.6f %lld\n", kv.second.first, (long long)kv.second.second);
    out += kv.first + line;
  }
  if (buf && cap > 0) {
    const size_t k = std::min<size_t>(out.size(), (size_t)cap - 1);
    std::memcpy(buf, out.data(), k);
    buf[k] = 0;
  }
  return (int32_t)out.size();
}

int32_t kg_plan_export(const kg_element* el, int32_t ne, int32_t n, int32_t kind, int32_t ordering,
                       const double* params, double s_rep, int32_t* buf, int32_t cap, int32_t* layout,
                       int32_t layout_cap) {
  // kind 3: the compact plan of the LDS-resident linear transient kernel
  const bool lin = kind == 3;
  if (!el || ne <= 0 || n <= 0 || kind < 0 || kind > 3 || ordering < 0 || ordering > ORD_BALANCED)
    return -KG_INVALID_ARGUMENT;
  PatLayout L;
  std::string err;
  std::vector<int32_t> b =
      build_pattern(el, ne, n, lin ? AK_TRAN : (AnalysisKind)kind, (Ordering)ordering, params, s_rep, L, err);
  if (layout) {
    int32_t m = std::min<int32_t>(layout_cap, (int32_t)(sizeof(PatLayout) / sizeof(int32_t)));
    std::memcpy(layout, &L, sizeof(int32_t) * m);
  }
  if (L.status != KG_OK) return -L.status;
  if (lin) {
    std::vector<uint16_t> w;
    if (!build_lin_plan(L, b, w)) return -KG_UNEXPECTED;
    if (buf) std::memcpy(buf, w.data(), sizeof(uint16_t) * std::min<size_t>(2 * (size_t)cap, w.size()));
    return (int32_t)(w.size() / 2);
  }
  if (buf) std::memcpy(buf, b.data(), sizeof(int32_t) * std::min<size_t>(cap, b.size()));
  return (int32_t)b.size();
}

int32_t kg_jit_ac_source(const kg_element* el, int32_t ne, int32_t n, const double* params, double s_rep,
                         int32_t precision, int32_t refine_iters, char* buf, int32_t cap) {
  if (!el || ne <= 0 || n <= 0) return -KG_INVALID_ARGUMENT;
  PatLayout L;
  std::string err, src, why;
  std::vector<int32_t> b = build_pattern(el, ne, n, AK_AC, ORD_MINDEG, params, s_rep, L, err);
  if (L.status != KG_OK) return -L.status;
  // refine_iters < 0 selects the mixed precision form with the split solution
  if (!jit_ac_source(L, b.data(), precision == KG_MIXED, refine_iters < 0, src, why)) {
    if (buf && cap > 0) std::snprintf(buf, cap, "%s", why.c_str());
    return -KG_UNEXPECTED;
  }
  if (buf && cap > 0) std::snprintf(buf, cap, "%s", src.c_str());
  return (int32_t)src.size();
}

int32_t kg_jit_dc_source(const kg_element* el, int32_t ne, int32_t n, const double* params, char* buf, int32_t cap) {
  if (!el || ne <= 0 || n <= 0) return -KG_INVALID_ARGUMENT;
  PatLayout L;
  std::string err, src, why;
  std::vector<int32_t> b = build_pattern(el, ne, n, AK_DC, ORD_MINDEG, params, 0.0, L, err);
  if (L.status != KG_OK) return -L.status;
  if (!jit_dc_source(L, b.data(), src, why)) {
    if (buf && cap > 0) std::snprintf(buf, cap, "%s", why.c_str());
    return -KG_UNEXPECTED;
  }
  if (buf && cap > 0) std::snprintf(buf, cap, "%s", src.c_str());
  return (int32_t)src.size();
}

int32_t kg_jit_tran_source(const kg_element* el, int32_t ne, int32_t n, const double* params, double h,
                           int32_t method, char* buf, int32_t cap) {
  if (!el || ne <= 0 || n <= 0 || !(h > 0)) return -KG_INVALID_ARGUMENT;
  PatLayout L;
  std::string err, src, why;
  std::vector<int32_t> b =
      build_pattern(el, ne, n, AK_TRAN, ORD_MINDEG, params, (method == KG_TRAP ? 2.0 : 1.0) / h, L, err);
  if (L.status != KG_OK) return -L.status;
  if (!jit_tran_source(L, b.data(), src, why)) {
    if (buf && cap > 0) std::snprintf(buf, cap, "%s", why.c_str());
    return -KG_UNEXPECTED;
  }
  if (buf && cap > 0) std::snprintf(buf, cap, "%s", src.c_str());
  return (int32_t)src.size();
}

int32_t kg_jit_lte_source(const kg_element* el, int32_t ne, int32_t n, const double* params, double h0, char* buf,
                          int32_t cap) {
  if (!el || ne <= 0 || n <= 0 || !(h0 > 0)) return -KG_INVALID_ARGUMENT;
  PatLayout L;
  std::string err, src, why;
  std::vector<int32_t> b = build_pattern(el, ne, n, AK_TRAN, ORD_MINDEG, params, 1.0 / h0, L, err);
  if (L.status != KG_OK) return -L.status;
  if (!jit_lte_source(L, b.data(), src, why)) {
    if (buf && cap > 0) std::snprintf(buf, cap, "%s", why.c_str());
    return -KG_UNEXPECTED;
  }
  if (buf && cap > 0) std::snprintf(buf, cap, "%s", src.c_str());
  return (int32_t)src.size();
}

int32_t kg_transient_lte(kg_ctx* c, const kg_solver_config* cfg, double tstop, double h0, double hmin, double hmax,
                         double lte_reltol, double lte_abstol, int32_t max_points, double* t, double* x,
                         int32_t* n_points, int32_t* newton_total, int32_t* rejected, int32_t* status, int32_t mem) {
  if (!c || !cfg || c->n_inst == 0 || !(tstop > 0) || !(h0 > 0) || !(hmin > 0) || !(hmax > 0) || max_points < 2 ||
      !(lte_reltol >= 0) || !(lte_abstol > 0))
    return set_err(c, KG_INVALID_ARGUMENT, "bad arguments");
  c->err.clear();
  HIPCHK(c, hipSetDevice(c->device));
  kg_solver_config ocfg = *cfg;
  ocfg.threads_per_system = 0;
  ocfg.compat = 0;
  double* d_xop;
  int64_t* d_xoff;
  int32_t *d_st, *d_it;
  int32_t r = run_dc(c, &ocfg, -1, 0.0, 0.0, 1, &d_xop, &d_xoff, &d_st, &d_it);
  if (r) return r;
  Plan P0;
  double s_tr = 1.0 / h0;
  if ((r = prepare(c, AK_TRAN, ORD_NESTED, s_tr, P0))) return r;
  auto off = out_offsets(c, max_points);
  int64_t* d_off;
  if ((r = upload_offsets(c, "lte_off", off, &d_off))) return r;
  hipError_t e;
  size_t ni = c->n_inst;
  double* d_t = (double*)ws_get(c, "lte_t", sizeof(double) * ni * max_points, &e);
  double* d_x = (double*)ws_get(c, "lte_x", sizeof(double) * off.back(), &e);
  int32_t* d_np = (int32_t*)ws_get(c, "lte_np", sizeof(int32_t) * ni, &e);
  int32_t* d_nt = (int32_t*)ws_get(c, "lte_nt", sizeof(int32_t) * ni, &e);
  int32_t* d_rj = (int32_t*)ws_get(c, "lte_rj", sizeof(int32_t) * ni, &e);
  int32_t* d_st0 = (int32_t*)ws_get(c, "lte_st0", sizeof(int32_t) * ni, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed");
  // every run (autotuner trials included) starts from the operating point's statuses
  HIPCHK(c, hipMemcpyAsync(d_st0, d_st, sizeof(int32_t) * ni, hipMemcpyDeviceToDevice, c->stream));
  NewtonArgs na = newton_args(cfg);
  na.compat = 0;
  LteArgs la{tstop, h0, hmin, hmax, lte_reltol, lte_abstol, max_points};
  // A mapping candidate: 1..256 threads per system (the generic kernel, plan
  // interpreted from LDS) or 2001 (one lane per instance, code specialised
  // to the shared topology at run time).
  struct LteLaunch {
    bool lane = false;
    int tps = 0, pi = 0;
    int64_t gchunk = 0;  // systems larger than LDS: systems per launch
    Plan P;
    Lay lay{};
    Launch L;
    JitFn fn;
  };
  auto setup = [&](int cand, LteLaunch& A) -> int32_t {
    A.lane = cand == 2001;
    if (A.lane) {
      if (!c->shared) return set_err(c, KG_UNEXPECTED, "lane kernel needs a shared topology");
      if (any_junction_charge(c)) return set_err(c, KG_UNEXPECTED, "lane kernel: junction charge is not generated");
      int32_t rr = prepare(c, AK_TRAN, ORD_MINDEG, s_tr, A.P);
      if (rr) return rr;
      if (A.P.group_status[0] != KG_OK) return set_err(c, KG_UNEXPECTED, "lane kernel: no valid plan");
      const int pidx = c->pat_index[std::make_tuple(0, (int)AK_TRAN, (int)ORD_MINDEG, s_tr)];
      const auto key = std::make_tuple(pidx, 100, 0);
      auto it = c->lane_fn.find(key);
      if (it == c->lane_fn.end()) {
        const PatEntry& pe = c->pats[pidx];
        std::string src, why, log;
        if (!jit_lte_source(pe.lay, pe.host.data(), src, why)) return set_err(c, KG_UNEXPECTED, "lane kernel: " + why);
        if (jit_load(c->device, src, jit_lte_name(), A.fn, log) != hipSuccess)
          return set_err(c, KG_UNEXPECTED, "lane kernel compilation failed: " + log);
        c->lane_fn[key] = A.fn;
      } else {
        A.fn = it->second;
      }
      A.L.grid = dim3((unsigned)((ni + 63) / 64));
      A.L.block = dim3(64);
      A.L.lds = 0;
      return KG_OK;
    }
    A.tps = cand;
    if (!valid_tps(A.tps)) return set_err(c, KG_INVALID_ARGUMENT, "threads_per_system must be 1, 4, 16, 64, 256 or 2001");
    int32_t rr = prepare(c, AK_TRAN, order_for(A.tps), s_tr, A.P);
    if (rr) return rr;
    A.lay = Lay{A.P.nnz_max, A.P.n_max, A.P.ne_max, 1, 2, 1, 8, 8, 0, 1};
    // stage the plan and each system's element records in LDS when they fit
    // (the whole step loop then reads them there); else the plan alone; else
    // neither; else (systems larger than LDS) the values in a global workspace
    A.pi = A.P.total_max;
    if (configure(c, A.tps, A.lay, c->n_inst, A.L, lay_al(4 * (size_t)A.pi)) != KG_OK) {
      A.lay.stage_prm = 0;
      if (configure(c, A.tps, A.lay, c->n_inst, A.L, lay_al(4 * (size_t)A.pi)) != KG_OK) {
        A.pi = 0;
        if ((rr = configure(c, A.tps, A.lay, c->n_inst, A.L))) {
          if (A.tps < 64) return rr;
          c->err.clear();
          if ((rr = configure_global(c, A.tps, A.lay, (int64_t)ni, A.L, "lte_gws", &A.gchunk))) return rr;
        }
      }
    }
    c->err.clear();
    return KG_OK;
  };
  auto run = [&](LteLaunch& A) -> hipError_t {
    hipError_t he = hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * ni, hipMemcpyDeviceToDevice, c->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(d_nt, d_it, sizeof(int32_t) * ni, hipMemcpyDeviceToDevice, c->stream);
    if (he == hipSuccess) he = hipMemsetAsync(d_t, 0, sizeof(double) * ni * max_points, c->stream);
    if (he != hipSuccess) return he;
    if (A.lane) {
      BatchDev B = batch_dev(c, A.P);
      const double* prm = B.params;
      const double* eb = B.ebase;
      int nia = (int)ni, mp = max_points, mi = na.maxiter;
      double a_tstop = tstop, a_h0 = h0, a_hmin = hmin, a_hmax = hmax, a_rel = lte_reltol, a_abs = lte_abstol;
      double a_reltol = na.reltol, a_itol = na.iabstol, a_vtol = na.vabstol;
      void* args[] = {&prm,    &eb,   &d_xop, &d_xoff, &d_t,   &d_x,     &d_off,  &d_np,   &d_nt,
                      &d_rj,   &d_st, &nia,   &a_tstop, &a_h0, &a_hmin, &a_hmax, &a_rel, &a_abs,
                      &mp,     &a_reltol, &a_itol, &a_vtol, &mi};
      return hipModuleLaunchKernel(A.fn.fn, A.L.grid.x, 1, 1, 64, 1, 1, 0, c->stream, args, nullptr);
    }
    BatchDev B = batch_dev(c, A.P);
    return chunked(A.lay, (int64_t)ni, A.gchunk, A.L, [&](Lay ly, dim3 g) {
      return run_lte(A.tps, g, A.L.block, A.L.lds, c->stream, B, 0, na, la, d_xop, d_xoff, d_t, d_x, d_off, d_np,
                     d_nt, d_rj, d_st, ly, A.pi);
    });
  };
  int cand = cfg->threads_per_system;
  if (cand == 0) {
    std::vector<int> cands{default_tps(c, P0.n_max)};
    if (c->shared) cands.push_back(2001);
    r = autotune(c, tune_key(c, "lte", KG_FP64, P0, (int64_t)ni), cands, [&](int tc, float* ms) -> int32_t {
      LteLaunch A;
      int32_t rr = setup(tc, A);
      if (rr) return rr;
      return time_launch(c, [&] { return run(A); }, ms);
    }, &cand);
    if (r) return r;
  }
  LteLaunch A;
  if ((r = setup(cand, A))) return r;
  r = timed(c, A.lane ? jit_lte_name() : "k_tran_lte", [&] { return run(A); });
  if (r) return r;
  return finish_status(c, d_st, status, mem,
                       {{t, d_t, sizeof(double) * ni * max_points, true}, {x, d_x, sizeof(double) * off.back(), true},
                        {n_points, d_np, sizeof(int32_t) * ni, true}, {newton_total, d_nt, sizeof(int32_t) * ni, true},
                        {rejected, d_rj, sizeof(int32_t) * ni, true}});
}

int32_t kg_adjoint_forward(kg_ctx* c, const kg_solver_config* cfg, double h, int32_t nsteps, int32_t method,
                           int32_t precision, const double* cvec, int32_t K, double* J, int32_t* status, int32_t mem) {
  if (!c || !cfg || c->n_inst == 0 || !(h > 0) || nsteps < 1 || K < 1 || !cvec ||
      (method != KG_BE && method != KG_TRAP))
    return set_err(c, KG_INVALID_ARGUMENT, "bad arguments (adjoint needs KG_BE or KG_TRAP, nsteps >= 1)");
  if (!c->shared) return set_err(c, KG_INVALID_ARGUMENT, "adjoint: batches with a shared topology only");
  c->err.clear();
  c->adj.ready = false;
  HIPCHK(c, hipSetDevice(c->device));
  kg_solver_config ocfg = *cfg;
  ocfg.threads_per_system = 0;
  ocfg.compat = 0;
  double* d_xop;
  int64_t* d_xoff;
  int32_t *d_st, *d_it;
  int32_t r = run_dc(c, &ocfg, -1, 0.0, 0.0, 1, &d_xop, &d_xoff, &d_st, &d_it);
  if (r) return r;
  Plan P0, P;
  double s_tr = (method == KG_TRAP ? 2.0 : 1.0) / h;
  if ((r = prepare(c, AK_TRAN, ORD_NESTED, s_tr, P0))) return r;
  if (P0.any_nonlinear) return set_err(c, KG_INVALID_ARGUMENT, "adjoint: linear circuits only");
  bool mixed = precision == KG_MIXED;
  int n = c->nunk[0];
  size_t ni = c->n_inst;
  int ncp = nsteps / K + 1, cpw = P0.n_max + P0.ne_max;
  hipError_t e;
  double* d_ck = (double*)ws_get(c, "adj_ckpt", sizeof(double) * ni * ncp * cpw, &e);
  double* d_c = (double*)ws_get(c, "adj_cvec", sizeof(double) * n, &e);
  double* d_J = (double*)ws_get(c, "adj_J", sizeof(double) * ni, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (checkpoints)");
  HIPCHK(c, hipMemcpyAsync(d_c, cvec, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  {  // keep the operating points and their offsets for the backward pass
    double* a_x = (double*)ws_get(c, "adj_xop", sizeof(double) * ni * n, &e);
    int64_t* a_o = (int64_t*)ws_get(c, "adj_xoff", sizeof(int64_t) * (ni + 1), &e);
    if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (adjoint)");
    HIPCHK(c, hipMemcpyAsync(a_x, d_xop, sizeof(double) * ni * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(a_o, d_xoff, sizeof(int64_t) * (ni + 1), hipMemcpyDeviceToDevice, c->stream));
  }
  AdjArgs aa{h, method == KG_TRAP ? 2.0 : 1.0, nsteps, method, cfg->refine_iters > 0 ? cfg->refine_iters : 2, K,
             method == KG_TRAP ? 1 : 0};
  // LDS-resident FP64 adjoint (k_adj_lin_fwd / _bwd): mapping 3000 + t
  const size_t ladj = mixed ? 0 : lin_adj_need(c, s_tr);
  auto lin_feasible = [&](int cand) {
    return ladj > 0 && ladj <= LDS_MAX && adj_lin_rows(cand - 3000, P0.n_max) > 0;
  };
  int32_t* d_st0 = nullptr;
  auto run_lin_fwd = [&](int cand) -> hipError_t {
    int nsrc_max = 0;
    int32_t rr = KG_OK;
    double* srcv = adj_srcv(c, s_tr, nsteps, &nsrc_max, &rr);
    if (rr) return hipErrorOutOfMemory;
    TranArgs ta{h, aa.alpha, nsteps, method, 0};
    return run_adj_lin_fwd(cand - 3000, adj_lin_rows(cand - 3000, P0.n_max), dim3((unsigned)ni), ladj, c->stream,
                           batch_dev(c, P0), 0, ta, K, d_c, d_xop, d_xoff, d_ck, (int64_t)ncp * cpw, cpw, d_J, d_st,
                           srcv, nsrc_max);
  };
  int tps = cfg->threads_per_system;
  if (tps == 0) {  // autotune the forward pass over the feasible mappings, in the requested precision
    d_st0 = (int32_t*)ws_get(c, "adj_status0", sizeof(int32_t) * ni, &e);
    if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (adjoint)");
    HIPCHK(c, hipMemcpyAsync(d_st0, d_st, sizeof(int32_t) * ni, hipMemcpyDeviceToDevice, c->stream));
    const int tdef = default_tps(c, P0.n_max);
    std::vector<int> cands{tdef};
    for (int t : {16, 64, 256})
      if (t != tdef) cands.push_back(t);
    if (!mixed && ladj > 0) cands.insert(cands.end(), {3064, 3256, 3512});
    int32_t rr = autotune(c, tune_key(c, "adjf", precision, P0, c->n_inst), cands,
                          [&](int cand, float* ms) -> int32_t {
      if (is_lin_cand(cand) && !lin_feasible(cand)) return set_err(c, KG_UNEXPECTED, "infeasible");
      HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * ni, hipMemcpyDeviceToDevice, c->stream));
      if (is_lin_cand(cand)) return time_launch(c, [&] { return run_lin_fwd(cand); }, ms);
      Plan PP;
      int32_t r2 = prepare(c, AK_TRAN, order_for(cand), s_tr, PP);
      if (r2) return r2;
      Lay ly{PP.nnz_max, PP.n_max, PP.ne_max, 0, 5, 1, mixed ? 4 : 8, 8, mixed ? 1 : 0, 0};
      Launch LL;
      int pi2 = (cand == 64 || cand == 256) ? PP.total_max : 0;
      if (pi2 && configure(c, cand, ly, c->n_inst, LL, lay_al(4 * (size_t)pi2)) != KG_OK) pi2 = 0;
      if (!pi2 && (r2 = configure(c, cand, ly, c->n_inst, LL))) return r2;
      BatchDev BB = batch_dev(c, PP);
      return time_launch(c, [&] {
        return run_adj_fwd(cand, mixed ? 1 : 0, LL.grid, LL.block, LL.lds, c->stream, BB, 0, aa, d_c, d_xop, d_xoff,
                           d_ck, (int64_t)ncp * cpw, cpw, d_J, d_st, ly, pi2);
      }, ms);
    }, &tps);
    if (rr) return rr;
    HIPCHK(c, hipMemcpyAsync(d_st, d_st0, sizeof(int32_t) * ni, hipMemcpyDeviceToDevice, c->stream));
  }
  if (tps == 0) tps = default_tps(c, P0.n_max);
  if (is_lin_cand(tps)) {
    if (!lin_feasible(tps))
      return set_err(c, KG_INVALID_ARGUMENT, "adjoint: the LDS-resident mapping needs FP64 and a fitting system");
    r = timed(c, "k_adj_lin_fwd", [&] { return run_lin_fwd(tps); });
    if (r) return r;
    r = finish_status(c, d_st, status, mem, {{J, d_J, sizeof(double) * ni, true}});
    c->adj.ready = true;
    c->adj.aa = aa;
    c->adj.tps = tps;
    c->adj.mixed = 0;
    c->adj.ncp = ncp;
    c->adj.cpw = cpw;
    return r;
  }
  if (!valid_tps(tps)) return set_err(c, KG_INVALID_ARGUMENT, "threads_per_system must be 1, 4, 16, 64 or 256");
  if ((r = prepare(c, AK_TRAN, order_for(tps), s_tr, P))) return r;
  Lay lay{P.nnz_max, P.n_max, P.ne_max, 0, 5, 1, mixed ? 4 : 8, 8, mixed ? 1 : 0, 0};
  Launch L;
  // stage the plan in LDS for the whole step loop when it fits
  int pi = (tps == 64 || tps == 256) ? P.total_max : 0;
  if (pi && configure(c, tps, lay, c->n_inst, L, lay_al(4 * (size_t)pi)) != KG_OK) pi = 0;
  if (!pi && (r = configure(c, tps, lay, c->n_inst, L))) return r;
  c->err.clear();
  BatchDev B = batch_dev(c, P);
  r = timed(c, mixed ? "k_adj_fwd<float,double>" : "k_adj_fwd<double,double>", [&] {
    return run_adj_fwd(tps, mixed ? 1 : 0, L.grid, L.block, L.lds, c->stream, B, 0, aa, d_c, d_xop, d_xoff, d_ck,
                       (int64_t)ncp * cpw, cpw, d_J, d_st, lay, pi);
  });
  if (r) return r;
  r = finish_status(c, d_st, status, mem, {{J, d_J, sizeof(double) * ni, true}});
  c->adj.ready = true;
  c->adj.aa = aa;
  c->adj.tps = tps;
  c->adj.mixed = mixed ? 1 : 0;
  c->adj.ncp = ncp;
  c->adj.cpw = cpw;
  c->adj.lay = lay;
  return r;
}

int32_t kg_adjoint_backward(kg_ctx* c, double* grad, int32_t* status, int32_t mem) {
  if (!c || !c->adj.ready) return set_err(c, KG_INVALID_ARGUMENT, "kg_adjoint_backward: no forward pass");
  c->err.clear();
  HIPCHK(c, hipSetDevice(c->device));
  const AdjArgs& aa = c->adj.aa;
  Plan P;
  double s_tr = aa.alpha / aa.h;
  int32_t r;
  if (is_lin_cand(c->adj.tps)) return adjoint_backward_lin(c, grad, status, mem);
  if ((r = prepare(c, AK_TRAN, order_for(c->adj.tps), s_tr, P))) return r;
  Lay lay = c->adj.lay;
  Launch L;
  int pi = (c->adj.tps == 64 || c->adj.tps == 256) ? P.total_max : 0;
  if (pi && configure(c, c->adj.tps, lay, c->n_inst, L, lay_al(4 * (size_t)pi)) != KG_OK) pi = 0;
  if (!pi && (r = configure(c, c->adj.tps, lay, c->n_inst, L))) return r;
  c->err.clear();
  hipError_t e;
  size_t ni = c->n_inst;
  int ne = (int)(c->eoff[1] - c->eoff[0]);
  int64_t seg_stride = (int64_t)(aa.K + 1) * P.n_max;
  double* d_ck = (double*)ws_get(c, "adj_ckpt", 8, &e);
  double* d_c = (double*)ws_get(c, "adj_cvec", 8, &e);
  double* d_seg = (double*)ws_get(c, "adj_seg", sizeof(double) * ni * seg_stride, &e);
  double* d_g = (double*)ws_get(c, "adj_grad", sizeof(double) * ni * ne, &e);
  double* d_xb = (double*)ws_get(c, "adj_xbar0", sizeof(double) * ni * P.n_max, &e);
  double* d_xop = (double*)ws_get(c, "adj_xop", 8, &e);
  int64_t* d_xoff = (int64_t*)ws_get(c, "adj_xoff", 8, &e);
  int32_t* d_st = (int32_t*)ws_get(c, "adj_status", sizeof(int32_t) * ni, &e);
  if (e != hipSuccess) return set_err(c, KG_HIP_ERROR, "workspace allocation failed (adjoint)");
  HIPCHK(c, hipMemsetAsync(d_st, 0, sizeof(int32_t) * ni, c->stream));
  BatchDev B = batch_dev(c, P);
  r = timed(c, c->adj.mixed ? "k_adj_bwd<float,double>" : "k_adj_bwd<double,double>", [&] {
    return run_adj_bwd(c->adj.tps, c->adj.mixed, L.grid, L.block, L.lds, c->stream, B, 0, aa, d_c, d_ck,
                       (int64_t)c->adj.ncp * c->adj.cpw, c->adj.cpw, d_seg, seg_stride, d_g, ne, d_xb, d_xoff, d_st,
                       lay, pi);
  });
  if (r) return r;
  // operating-point part (DC plan, FP64)
  Plan PD;
  int tpd = default_tps(c, P.n_max);
  if ((r = prepare(c, AK_DC, order_for(tpd), 0.0, PD))) return r;
  Lay ld{PD.nnz_max, PD.n_max, PD.ne_max, 0, 0, 0, 8, 8, 0, 0};
  Launch LD;
  if ((r = configure(c, tpd, ld, c->n_inst, LD))) return r;
  BatchDev BD = batch_dev(c, PD);
  r = timed(c, "k_adj_op", [&] {
    return run_adj_op(tpd, LD.grid, LD.block, LD.lds, c->stream, BD, 0, d_xb, d_xop, d_xoff, d_g, ne, d_st, ld);
  });
  if (r) return r;
  return finish_status(c, d_st, status, mem, {{grad, d_g, sizeof(double) * ni * ne, true}});
}

int32_t kg_adjoint_transient(kg_ctx* c, const kg_solver_config* cfg, double h, int32_t nsteps, int32_t method,
                             int32_t precision, const double* cvec, int32_t K, double* J, double* grad,
                             int32_t* status, int32_t mem) {
  int32_t r = kg_adjoint_forward(c, cfg, h, nsteps, method, precision, cvec, K, J, status, mem);
  if (r != KG_OK) return r;
  return kg_adjoint_backward(c, grad, status, mem);
}

}  // extern "C"
