"""CPU checks of the run-time specialised lane-per-system AC kernel.

The generated HIP source (kg_jit_ac_source) is compiled twice without a GPU:
  * for gfx950 with hipcc (device only), checking that it builds and, for
    every form on the 50-section RC ladder, that it needs no scratch memory
    (the whole system lives in registers);
  * for the host with g++, through a small header that maps the handful of
    HIP names it uses onto host C++, and run lane by lane over a batch: its
    results are compared with the oracle's AC solves, so the specialised
    factorisation (scaled upper factor, fused forward substitution) is
    checked before any GPU run.
"""
import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import active_netlist, read_circuit, time_aligned

HIPCC = "/opt/rocm/bin/hipcc"

HOST_SHIM = r"""
#include <cmath>
#include <cstring>
#define __device__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(x)
struct kg_dim3 { unsigned x; };
static kg_dim3 threadIdx, blockIdx, gridDim;
static inline int kg_frexp_exp(double a) { int e = 0; std::frexp(a, &e); return e; }
#define __builtin_amdgcn_frexp_exp(a) kg_frexp_exp(a)
#define __builtin_amdgcn_ldexp(a, e) std::ldexp((a), (e))
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
// the hardware reciprocal estimate: off by up to 2^-24 relative (measured on
// gfx950: at most 4.6e-8 = 2^-24.4); the host takes the full 2^-24, with a
// sign that varies with the operand, so the host runs see the estimate's
// worst case
static inline double kg_rcp(double a) {
  const double r = 1.0 / a;
  unsigned long long b;
  std::memcpy(&b, &a, 8);
  return r * (1.0 + (((b >> 29) & 1) ? 0x1p-24 : -0x1p-24));
}
#define __builtin_amdgcn_rcp(a) kg_rcp(a)
// one lane at a time: a wave vote is the lane's own predicate
#define __builtin_amdgcn_ballot_w64(p) ((p) ? ~0ull : 0ull)
#define __builtin_amdgcn_read_exec() (~0ull)
// a shuffle sees the lane's own value and every lane is its own segment: each
// publishes its own figure, whose maximum is what the wave reduction publishes
template <class T> static inline T __shfl_down(T v, unsigned, int) { return v; }
static inline int __shfl_up(int v, unsigned, int) { return v + 1; }
static inline int atomicMax(int* p, int v) { int o = *p; if (v > o) *p = v; return o; }
static inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  unsigned long long o = *p; if (v > o) *p = v; return o; }
static inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
static inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  unsigned long long o = *p; *p = o + v; return o; }
using std::fmax; using std::sqrt;
// raw buffer loads of the mixed precision kernels, bounds-checked on the host
struct kg_rsrc { const char* p; long long n; };
typedef kg_rsrc __amdgpu_buffer_rsrc_t;
static int kg_oob = 0;
static inline unsigned long long kg_bload(kg_rsrc r, int v, int s) {
  if (v < 0 || s < 0 || (long long)v + s + 8 > r.n) { kg_oob++; return 0; }
  unsigned long long x; std::memcpy(&x, r.p + v + s, 8); return x; }
#define __builtin_amdgcn_make_buffer_rsrc(p, st, n, f) (kg_rsrc{(const char*)(p), (long long)(n)})
#define __builtin_amdgcn_raw_buffer_load_b64(r, v, s, a) kg_bload((r), (v), (s))
#define __builtin_amdgcn_readfirstlane(x) (x)
"""

HOST_DRIVER = r"""
extern "C" void run_host(const double* prm, const double* eb, const double* fr, int nf, long long nsys,
                         double* out, int* st, double* err, int refine, double stop, int* oob,
                         unsigned long long* fixq, long long fixcap) {
  gridDim.x = (unsigned)((nsys + 63) / 64);
  for (unsigned b = 0; b < gridDim.x; b++)
    for (unsigned t = 0; t < 64; t++) {
      blockIdx.x = b;
      threadIdx.x = t;
      KERNEL(prm, eb, fr, nf, nsys, out, st, err, refine, stop, fixq, fixcap);
    }
  *oob = kg_oob;
}
"""


def source(c, precision, refine=2, s_rep=2 * np.pi * 1e3):
    """Generated source; refine < 0 selects the split-solution mixed form."""
    from krets_amd._lib import lib
    buf = ctypes.create_string_buffer(1 << 23)
    el = np.ascontiguousarray(c.elements)
    pr = np.ascontiguousarray(c.params)
    r = lib().kg_jit_ac_source(el.ctypes.data, len(el), c.n, pr.ctypes.data, s_rep, precision, refine, buf, len(buf))
    return r, buf.value.decode()


def ebase(elements, params):
    e = params[..., 0].copy()
    isr = (elements["type"] == 0) & (elements["g2"] == 0)
    with np.errstate(divide="ignore"):  # a zero resistor's conductance is infinite, as in the library
        e[..., isr] = 1.0 / e[..., isr]
    return e


def host_run(src, params, elements, freqs, n, refine=2, stop=0.0, queued=None):
    """Runs the generated kernel lane by lane on the host. queued (a list)
    receives the systems the accuracy guard queued for the fixup."""
    d = tempfile.mkdtemp()
    try:
        cpp = os.path.join(d, "k.cpp")
        so = os.path.join(d, "k.so")
        # the LDS array of the split-solution form: one host array (lanes run
        # one after another and index it by lane)
        src = src.replace("extern __shared__ W xs[];", "static W xs[64 * 128];")
        name = re.search(r"void __launch_bounds__\(64\) (\w+)\(", src).group(1)
        open(cpp, "w").write(HOST_SHIM + src + HOST_DRIVER.replace("KERNEL", name))
        subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", cpp, "-o", so],
                       check=True, capture_output=True)
        L = ctypes.CDLL(so)
        ni, nf = params.shape[0], len(freqs)
        prm = np.ascontiguousarray(params, np.float64)
        eb = ebase(elements, params)
        if "c64" in name:  # the mixed precision kernels read the table in blocks of 64 instances
            nb = (ni + 63) // 64
            t = np.zeros((nb * 64, eb.shape[1]))
            t[:ni] = eb
            eb = t.reshape(nb, 64, -1).transpose(0, 2, 1)
        eb = np.ascontiguousarray(eb, np.float64)
        fr = np.ascontiguousarray(freqs, np.float64)
        out = np.zeros(ni * nf * n * 2)
        st = np.zeros(ni, np.int32)
        err = np.zeros(ni)
        P = ctypes.c_void_p
        L.run_host.argtypes = [P, P, P, ctypes.c_int, ctypes.c_longlong, P, P, P, ctypes.c_int, ctypes.c_double, P,
                               P, ctypes.c_longlong]
        oob = np.zeros(1, np.int32)
        fixq = np.zeros(1 + ni * nf, np.uint64)
        L.run_host(prm.ctypes.data, eb.ctypes.data, fr.ctypes.data, nf, ni * nf, out.ctypes.data, st.ctypes.data,
                   err.ctypes.data, refine, stop, oob.ctypes.data, fixq.ctypes.data, ni * nf)
        assert oob[0] == 0, "buffer load outside the instance group's block"
        if queued is not None:
            queued.extend(int(v) for v in fixq[1:1 + int(fixq[0])])
        x = out.reshape(ni, nf, n, 2)
        return x[..., 0] + 1j * x[..., 1], st, err
    finally:
        shutil.rmtree(d)


def nerr(x, r):
    d = np.linalg.norm(r)
    return np.linalg.norm(x - r) / (d if d > 0 else 1.0)


@pytest.mark.parametrize("xsplit", [False, True])
@pytest.mark.parametrize("precision,refine", [(0, 2), (1, 1), (1, 2), (1, 3)])
def test_lane_source_on_host_vs_oracle(oracle, precision, refine, xsplit):
    import krets_amd as kg
    from krets_amd import workloads as W
    if xsplit and precision == 0:
        pytest.skip("the FP64 kernel keeps no solution between passes")
    b = W.c2_rc_ladder_mc(3, sections=50, seed=21)
    c = kg.Circuit(b.netlists[0])
    r, src = source(c, precision, -1 if xsplit else refine)
    assert r > 0, src
    f = W.c2_frequencies(200)[::25]
    queued = []
    x, st, err = host_run(src, b.params, c.elements, f, c.n, refine, queued=queued)
    assert (st == 0).all()
    for i in range(3):
        xr = oracle.Circuit(b.netlists[i]).ac(f)
        e = max(nerr(x[i, j], xr[j]) for j in range(len(f)))
        assert e < (1e-12 if precision == 0 or refine >= 2 else 1e-7), (i, e)
    if precision == 0:  # the published backward-error bound c_m rho; the ladders' static LU has rho = 1
        import plan_emulator as pe
        plan = pe.Plan(c.elements, c.n, pe.AC, pe.MINDEG, c.params, 2 * np.pi * 1e3)
        m = int(np.max(np.diff(plan.row_ptr)))
        assert not queued and np.allclose(err, guard_bound_coeff(m), rtol=1e-12)
    elif refine >= 2:  # estimate: the last correction, scaled by the contraction seen
        assert not queued and (err > 0).all() and (err < 1e-10).all()
    else:  # one pass leaves estimates above the guard: those solves are queued for the FP64 fixup
        assert queued and (err < 1e-10).all()


@pytest.mark.parametrize("xsplit", [False, True])
def test_lane_source_mixed_across_instance_blocks_on_host(xsplit):
    """70 instances: a full block of 64 in the blocked value table and a
    partial one of 6. The host run bounds-checks every buffer load against
    its block; the mixed results agree with the FP64 kernel's."""
    import krets_amd as kg
    from krets_amd import workloads as W
    b = W.c2_rc_ladder_mc(70, sections=50, seed=31)
    c = kg.Circuit(b.netlists[0])
    f = W.c2_frequencies(200)[::40]
    _, src64 = source(c, 0)
    x64, _, _ = host_run(src64, b.params, c.elements, f, c.n)
    _, srcm = source(c, 1, -1 if xsplit else 2)
    xm, st, err = host_run(srcm, b.params, c.elements, f, c.n, refine=2)
    assert (st == 0).all()
    e = np.linalg.norm(xm - x64, axis=2) / np.linalg.norm(x64, axis=2)
    assert e.max() < 1e-9, e.max()


def test_lane_source_golden_ac_on_host():
    import json
    import krets_amd as kg
    from conftest import GOLDEN
    g = json.load(open(os.path.join(GOLDEN, "low_pass_ac_result.json")))
    cols = {k: np.array([float(v) for v in vals]) for k, vals in g["columns"].items()}
    c = kg.Circuit(read_circuit("low_pass_filter/ac/ac.cir"))
    r, src = source(c, 0)
    assert r > 0, src
    x, st, _ = host_run(src, c.params[None], c.elements, cols["frequency"], c.n)
    assert st[0] == 0
    for name in ("I(V1)", "V(in)", "V(out)"):
        v = x[0][:, c.names.index(name)]
        assert np.max(np.abs(np.abs(v) - cols[name + "_mag"]) / cols[name + "_mag"]) < 1e-12


def test_lane_source_singular_pivot_queued():
    # node a is reached only through zero-valued capacitors: its row and
    # column are exactly zero, so its pivot is zero (Decomposition)
    import krets_amd as kg
    c = kg.Circuit("V1 in 0 0 AC 1\nR1 in 0 1000\nC1 in a 0\nC2 a 0 0\n")
    r, src = source(c, 0)
    assert r > 0, src
    queued = []
    _, st, _ = host_run(src, c.params[None], c.elements, np.array([1e3]), c.n, queued=queued)
    # a failed static pivot is not yet a failed decomposition: the solve is queued for the fixup's
    # partial pivoting, which decides (tests/test_gpu_guard.py)
    assert queued == [0] and st[0] == 0


def test_lane_source_declines_diodes_and_large_systems():
    import krets_amd as kg
    from krets_amd import workloads as W
    c = kg.Circuit(read_circuit("rectifier/rectifier.cir"))
    r, why = source(c, 0)
    assert r < 0 and "diode" in why
    b = W.c2_rc_ladder_mc(1, sections=200, seed=1)
    r, why = source(kg.Circuit(b.netlists[0]), 0)
    assert r < 0 and "register" in why


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("precision,refine", [(0, 2), (1, 2), (1, -1)])  # -1: split solution
def test_lane_source_builds_for_gfx950(precision, refine):
    """Every generated form builds for gfx950 and keeps the 50-section
    ladder in registers: no scratch (spills) for FP64, the mixed form with
    an FP64 solution, and the split-solution mixed form with its adaptive
    refinement stop."""
    import krets_amd as kg
    from krets_amd import workloads as W
    b = W.c2_rc_ladder_mc(1, sections=50, seed=2)
    r, src = source(kg.Circuit(b.netlists[0]), precision, refine)
    assert r > 0, src
    d = tempfile.mkdtemp()
    try:
        p = os.path.join(d, "k.hip")
        open(p, "w").write("#include <hip/hip_runtime.h>\n" + src)
        out = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                              "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", p, "-o", p + ".o",
                              "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True)
        assert out.returncode == 0, out.stderr[-2000:]
        scratch = int(re.search(r"ScratchSize \[bytes/lane\]: (\d+)", out.stderr).group(1))
        assert scratch == 0
    finally:
        shutil.rmtree(d)


def test_lane_source_adaptive_stop_on_host(oracle):
    """The mixed kernel's adaptive stop (default target 1e-11 on the
    predicted error 200 * (|dx| / |x|)^2): on C2-like RC ladders lanes leave
    after one refinement pass where that suffices, every result stays within
    1e-9 of FP64, and the reported estimate bounds the true error."""
    import krets_amd as kg
    from krets_amd import workloads as W
    b = W.c2_rc_ladder_mc(16, sections=50, seed=2)
    c = kg.Circuit(b.netlists[0])
    f = W.c2_frequencies(200)
    _, src64 = source(c, 0)
    x64, _, _ = host_run(src64, b.params, c.elements, f, c.n)
    _, srcm = source(c, 1, -1)
    xa, st, err = host_run(srcm, b.params, c.elements, f, c.n, refine=2, stop=1e-11)
    x1, _, _ = host_run(srcm, b.params, c.elements, f, c.n, refine=1)
    x2, _, _ = host_run(srcm, b.params, c.elements, f, c.n, refine=2)
    assert (st == 0).all()
    e = np.linalg.norm(xa - x64, axis=2) / np.linalg.norm(x64, axis=2)
    assert e.max() < 1e-9, e.max()
    stopped = np.all(xa == x1, axis=2)
    assert 0.2 < stopped.mean() < 0.9, stopped.mean()  # some, not all, leave early
    assert np.array_equal(xa[~stopped], x2[~stopped])
    assert (err >= e.max(axis=1)).all() and (err <= 1e-9).all()


@pytest.mark.parametrize("precision,refine", [(0, 2), (1, 2), (1, -1)])
def test_lane_source_controlled_sources_on_host(oracle, precision, refine):
    """The generated kernel stamps E, G, F and H gains (the active filter of
    every controlled source, three instances with different values) and
    matches the oracle's AC solves."""
    import krets_amd as kg
    from conftest import active_netlist
    nets = [active_netlist(), active_netlist(15e3, 18e3, 2.2e-9, 0.8e-9, 1.5, 0.5, 3e-3, 2e3),
            active_netlist(8e3, 30e3, 6.8e-9, 1.2e-9, 0.9, 4.0, 0.5e-3, 100.0)]
    cs = [kg.Circuit(t) for t in nets]
    c = cs[0]
    params = np.stack([x.params for x in cs])
    r, src = source(c, precision, refine)
    assert r > 0, src
    f = np.array([10.0, 3e3, 2e4, 1e5, 1e6])
    x, st, _ = host_run(src, params, c.elements, f, c.n, refine=max(refine, 2), stop=0.0)
    assert (st == 0).all()
    for i, t in enumerate(nets):
        xr = oracle.Circuit(t).ac(f)
        e = max(nerr(x[i, j], xr[j]) for j in range(len(f)))
        assert e < (1e-12 if precision == 0 else 1e-9), (i, e)


# ------------------------------------------------- LTE transient lane kernel
LTE_SHIM = r"""
#include <cmath>
#define __device__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(x)
struct kg_dim3 { unsigned x; };
static kg_dim3 threadIdx, blockIdx, gridDim;
using std::exp; using std::log; using std::sin; using std::fmod; using std::sqrt;
using std::fmin; using std::fmax; using std::fabs;
"""

LTE_DRIVER = r"""
extern "C" void run_host(const double* prm, const double* eb, const double* xop, const long long* xop_off,
                         double* tout, double* xout, const long long* xout_off, int* npts, int* ntot, int* rej,
                         int* status, int ninst, double tstop, double h0, double hmin, double hmax, double lrel,
                         double labs, int maxp, double reltol, double itol, double vtol, int maxiter) {
  gridDim.x = (unsigned)((ninst + 63) / 64);
  for (unsigned b = 0; b < gridDim.x; b++)
    for (unsigned t = 0; t < 64; t++) {
      blockIdx.x = b;
      threadIdx.x = t;
      kg_lte_lane(prm, eb, xop, xop_off, tout, xout, xout_off, npts, ntot, rej, status, ninst, tstop, h0, hmin,
                  hmax, lrel, labs, maxp, reltol, itol, vtol, maxiter);
    }
}
"""


def lte_source(c, params, h0):
    from krets_amd._lib import lib
    buf = ctypes.create_string_buffer(1 << 22)
    el = np.ascontiguousarray(c.elements)
    pr = np.ascontiguousarray(params, np.float64)
    r = lib().kg_jit_lte_source(el.ctypes.data, len(el), c.n, pr.ctypes.data, h0, buf, len(buf))
    return r, buf.value.decode()


def lte_host_run(src, params, elements, x0, args, maxp, cfg):
    d = tempfile.mkdtemp()
    try:
        cpp, so = os.path.join(d, "k.cpp"), os.path.join(d, "k.so")
        open(cpp, "w").write(LTE_SHIM + src + LTE_DRIVER)
        subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", cpp, "-o", so], check=True,
                       capture_output=True)
        L = ctypes.CDLL(so)
        ni, n = x0.shape
        prm = np.ascontiguousarray(params, np.float64)
        eb = np.ascontiguousarray(ebase(elements, params), np.float64)
        xop = np.ascontiguousarray(x0, np.float64)
        xoff = np.arange(ni, dtype=np.int64) * n
        tout = np.zeros(ni * maxp)
        xout = np.zeros(ni * maxp * n)
        ooff = np.arange(ni, dtype=np.int64) * maxp * n
        npts, ntot, rej, st = (np.zeros(ni, np.int32) for _ in range(4))
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.run_host.argtypes = [P] * 11 + [I, D, D, D, D, D, D, I, D, D, D, I]
        ptrs = [a.ctypes.data for a in (prm, eb, xop, xoff, tout, xout, ooff, npts, ntot, rej, st)]
        L.run_host(*ptrs, ni, *args, maxp, *cfg)
        return tout.reshape(ni, maxp), xout.reshape(ni, maxp, n), npts, ntot, rej, st
    finally:
        shutil.rmtree(d)


def test_lte_lane_source_on_host_vs_oracle(oracle):
    """The generated LTE lane kernel, run lane by lane on the host, takes the
    oracle's accepted time grid, rejection count and Newton iterations on the
    half-wave rectifiers, and its states agree to 1e-9 once the oracle's are
    moved to the same times (time_aligned)."""
    import krets_amd as kg
    from krets_amd import workloads as W
    b = W.c4_rectifier(6, seed=4)
    c = kg.Circuit(b.netlists[0])
    args = (5e-3, 1e-6, 1e-14, 5e-5, 1e-3, 1e-6)
    r, src = lte_source(c, b.params[0], args[1])
    assert r > 0, src
    ocfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    refs = [oracle.Circuit(b.netlists[i]).tran_lte(*args, cfg=ocfg) for i in range(6)]
    x0 = np.stack([ref[1][0] for ref in refs])
    t, x, npts, nt, rej, st = lte_host_run(src, b.params, c.elements, x0, args, 4000, (1e-3, 1e-12, 1e-9, 300))
    assert (st == 0).all()
    op_iters = set()
    for i, (tr, xr, ntr, rejr) in enumerate(refs):
        assert npts[i] == len(tr) and rej[i] == rejr, (i, npts[i], len(tr), rej[i], rejr)
        op_iters.add(ntr - nt[i])  # the oracle's total includes its operating point
        assert np.max(np.abs(t[i][:npts[i]] - tr)) <= 1e-9 * 5e-3
        xa = time_aligned(tr, xr, t[i][:npts[i]])
        assert np.array_equal(x[i][0], xr[0])  # the operating point is the input
        assert max(nerr(x[i][j], xa[j]) for j in range(1, len(tr))) < 1e-9
    assert len(op_iters) == 1 and 1 <= op_iters.pop() <= 5


def test_lte_lane_source_builds_for_gfx950():
    """The generated LTE kernel builds for gfx950 without scratch on the
    rectifier and on the active filter of every controlled source."""
    import krets_amd as kg
    from conftest import active_netlist
    from krets_amd import workloads as W
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    b = W.c4_rectifier(1, seed=4)
    for c in (kg.Circuit(b.netlists[0]), kg.Circuit(active_netlist())):
        r, src = lte_source(c, c.params, 1e-6)
        assert r > 0, src
        d = tempfile.mkdtemp()
        try:
            p = os.path.join(d, "k.hip")
            open(p, "w").write("#include <hip/hip_runtime.h>\n" + src)
            out = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                  "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", p, "-o", p + ".o",
                                  "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True)
            assert out.returncode == 0, out.stderr[-2000:]
            assert int(re.search(r"ScratchSize \[bytes/lane\]: (\d+)", out.stderr).group(1)) == 0
        finally:
            shutil.rmtree(d)


# ------------------------------------------------- linear OP / DC lane kernel
DC_DRIVER = r"""
extern "C" void run_host(const double* prm, const double* eb, int ninst, int sweep_e, double start, double step,
                         int npts, double* out, const long long* off, int* status, int* iters, int compat,
                         double reltol, double itol, double vtol, int maxiter) {
  gridDim.x = (unsigned)((ninst + 63) / 64);
  for (unsigned b = 0; b < gridDim.x; b++)
    for (unsigned t = 0; t < 64; t++) {
      blockIdx.x = b;
      threadIdx.x = t;
      kg_dc_lane(prm, eb, ninst, sweep_e, start, step, npts, out, off, status, iters, compat, reltol, itol, vtol,
                 maxiter);
    }
}
"""


def dc_source(c, params):
    from krets_amd._lib import lib
    buf = ctypes.create_string_buffer(1 << 22)
    el = np.ascontiguousarray(c.elements)
    pr = np.ascontiguousarray(params, np.float64)
    r = lib().kg_jit_dc_source(el.ctypes.data, len(el), c.n, pr.ctypes.data, buf, len(buf))
    return r, buf.value.decode()


def dc_host_run(src, params, elements, n, sweep_e, start, step, npts, cfg=(1, 1e-3, 1e-12, 1e-6, 300)):
    d = tempfile.mkdtemp()
    try:
        cpp, so = os.path.join(d, "k.cpp"), os.path.join(d, "k.so")
        shim = LTE_SHIM.replace("#include <cmath>", "#include <cmath>\n#define asm(...)\n#define volatile")
        open(cpp, "w").write(shim + src + DC_DRIVER)
        subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", cpp, "-o", so], check=True,
                       capture_output=True)
        L = ctypes.CDLL(so)
        ni = params.shape[0]
        prm = np.ascontiguousarray(params, np.float64)
        eb = np.ascontiguousarray(ebase(elements, params), np.float64)
        out = np.zeros(ni * npts * n)
        off = np.arange(ni, dtype=np.int64) * npts * n
        st, it = np.zeros(ni, np.int32), np.zeros(ni * npts, np.int32)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.run_host.argtypes = [P, P, I, I, D, D, I, P, P, P, P, I, D, D, D, I]
        L.run_host(prm.ctypes.data, eb.ctypes.data, ni, sweep_e, start, step, npts, out.ctypes.data, off.ctypes.data,
                   st.ctypes.data, it.ctypes.data, *cfg)
        return out.reshape(ni, npts, n), st, it.reshape(ni, npts)
    finally:
        shutil.rmtree(d)


def test_dc_lane_source_on_host_vs_oracle(oracle):
    """The generated linear OP / DC-sweep kernel, run on the host: the C2 RC
    ladders' operating points, a DC sweep of the voltage divider and of the
    active filter of every controlled source, against the oracle (1e-12)."""
    import krets_amd as kg
    from conftest import active_netlist
    from krets_amd import workloads as W
    b = W.c2_rc_ladder_mc(4, sections=50, seed=3)
    c = kg.Circuit(b.netlists[0])
    r, src = dc_source(c, b.params[0])
    assert r > 0, src
    params = b.params.copy()
    params[:, 0, 0] = 1.0 + np.arange(4)  # V1's DC value (the sweep's is 0)
    x, st, it = dc_host_run(src, params, c.elements, c.n, -1, 0.0, 0.0, 1)
    assert (st == 0).all() and (it == 1).all()
    for i in range(4):
        oc = oracle.Circuit(b.netlists[i])
        oc.set_value(0, 1.0 + i)
        assert nerr(x[i][0], oc.op()[0]) < 1e-12
    for text in (read_circuit("voltage_divider/voltage_divider.cir"), active_netlist()):
        c = kg.Circuit(text)
        r, src = dc_source(c, c.params)
        assert r > 0, src
        x, st, _ = dc_host_run(src, c.params[None], c.elements, c.n, 0, -1.0, 0.25, 9)
        xr = oracle.Circuit(text).dc("V1", -1.0, 1.0, 0.25)
        assert st[0] == 0 and max(nerr(x[0][p], xr[p]) for p in range(9)) < 1e-12


def test_dc_lane_source_diodes_on_host_vs_oracle(oracle):
    """Circuits with diodes through krets' Newton semantics (empty first map,
    0.5 V default, the step key of sweeps): the diode I-V sweep and the
    rectifier's operating point, with the oracle's iteration counts."""
    import krets_amd as kg
    text = read_circuit("diode_iv_curve/diode_iv_curve.cir")
    c = kg.Circuit(text)
    r, src = dc_source(c, c.params)
    assert r > 0, src
    x, st, it = dc_host_run(src, c.params[None], c.elements, c.n, 0, 0.0, 0.25, 9)
    oc = oracle.Circuit(text)
    xr = oc.dc("V1", 0.0, 2.0, 0.25)
    assert st[0] == 0 and max(nerr(x[0][p], xr[p]) for p in range(9)) < 1e-9
    x, st, it = dc_host_run(src, c.params[None], c.elements, c.n, -1, 0.0, 0.0, 1)
    xo, ito = oc.op()
    assert st[0] == 0 and nerr(x[0][0], xo) < 1e-9 and it[0][0] == ito
    c = kg.Circuit(read_circuit("rectifier/rectifier.cir"))
    r, src = dc_source(c, c.params)
    assert r > 0, src
    x, st, it = dc_host_run(src, c.params[None], c.elements, c.n, -1, 0.0, 0.0, 1)
    xo, ito = oracle.Circuit(read_circuit("rectifier/rectifier.cir")).op()
    assert st[0] == 0 and nerr(x[0][0], xo) < 1e-9 and it[0][0] == ito


def test_dc_lane_source_flags_singular():
    import krets_amd as kg
    # a zero resistor's conductance is infinite: a non-finite pivot (Decomposition)
    c = kg.Circuit("V1 1 0 1\nR1 1 2 0\nR2 2 0 1k\n")
    r, src = dc_source(c, c.params)
    assert r > 0, src
    _, st, _ = dc_host_run(src, c.params[None], c.elements, c.n, -1, 0.0, 0.0, 1)
    assert st[0] == 2


def test_dc_lane_source_builds_for_gfx950():
    """The operating-point kernels of the 50-section ladder (linear) and of the
    rectifier (Newton) build for gfx950 without scratch."""
    import krets_amd as kg
    from krets_amd import workloads as W
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    b = W.c2_rc_ladder_mc(1, sections=50, seed=2)
    for c in (kg.Circuit(b.netlists[0]), kg.Circuit(read_circuit("rectifier/rectifier.cir"))):
        r, src = dc_source(c, c.params)
        assert r > 0, src
        d = tempfile.mkdtemp()
        try:
            p = os.path.join(d, "k.hip")
            open(p, "w").write("#include <hip/hip_runtime.h>\n" + src)
            out = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                  "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", p, "-o", p + ".o",
                                  "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True)
            assert out.returncode == 0, out.stderr[-2000:]
            assert int(re.search(r"ScratchSize \[bytes/lane\]: (\d+)", out.stderr).group(1)) == 0
        finally:
            shutil.rmtree(d)


# ------------------------------------------------- fixed-step transient lane kernel
TRAN_DRIVER = r"""
extern "C" void run_host(const double* prm, const double* eb, const double* xop, const long long* xop_off,
                         double* out, const long long* out_off, int* status, int* iters, double* err, int ninst,
                         double h, double alpha, int nsteps, int method, double reltol, double itol, double vtol,
                         int maxiter) {
  gridDim.x = (unsigned)((ninst + 63) / 64);
  for (unsigned b = 0; b < gridDim.x; b++)
    for (unsigned t = 0; t < 64; t++) {
      blockIdx.x = b;
      threadIdx.x = t;
      kg_tran_lane(prm, eb, xop, xop_off, out, out_off, status, iters, err, ninst, h, alpha, nsteps, method,
                   reltol, itol, vtol, maxiter);
    }
}
"""


def tran_source(c, params, h, method):
    from krets_amd._lib import lib
    buf = ctypes.create_string_buffer(1 << 22)
    el = np.ascontiguousarray(c.elements)
    pr = np.ascontiguousarray(params, np.float64)
    r = lib().kg_jit_tran_source(el.ctypes.data, len(el), c.n, pr.ctypes.data, h, method, buf, len(buf))
    return r, buf.value.decode()


def tran_host_run(src, params, elements, x0, h, nsteps, method, tol, with_err=False):
    d = tempfile.mkdtemp()
    try:
        cpp, so = os.path.join(d, "k.cpp"), os.path.join(d, "k.so")
        open(cpp, "w").write(LTE_SHIM + src + TRAN_DRIVER)
        subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", cpp, "-o", so], check=True,
                       capture_output=True)
        L = ctypes.CDLL(so)
        ni, n = x0.shape
        prm = np.ascontiguousarray(params, np.float64)
        eb = np.ascontiguousarray(ebase(elements, params), np.float64)
        xop = np.ascontiguousarray(x0, np.float64)
        xoff = np.arange(ni, dtype=np.int64) * n
        out = np.zeros(ni * (nsteps + 1) * n)
        ooff = np.arange(ni, dtype=np.int64) * (nsteps + 1) * n
        st, it, err = np.zeros(ni, np.int32), np.zeros(ni * (nsteps + 1), np.int32), np.zeros(ni)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.run_host.argtypes = [P] * 9 + [I, D, D, I, I, D, D, D, I]
        L.run_host(*[a.ctypes.data for a in (prm, eb, xop, xoff, out, ooff, st, it, err)], ni, h,
                   2.0 if method == 2 else 1.0, nsteps, method, *tol)
        res = (out.reshape(ni, nsteps + 1, n), st, it.reshape(ni, nsteps + 1))
        return res + (err,) if with_err else res
    finally:
        shutil.rmtree(d)


@pytest.mark.parametrize("method", [0, 1, 2])
def test_tran_lane_source_rectifier_on_host_vs_oracle(oracle, method):
    """The generated fixed-step kernel on the half-wave rectifiers at h = 1 us:
    krets' integrator (method 0, with its lagged sources, 0.5 V default and
    time key) and the clean BE / trapezoidal ones, against the oracle: the
    same Newton iterations at every step, states to 1e-9."""
    import krets_amd as kg
    from krets_amd import workloads as W
    b = W.c4_rectifier(4, seed=4)
    c = kg.Circuit(b.netlists[0])
    h, N = 1e-6, 600
    r, src = tran_source(c, b.params[0], h, method)
    assert r > 0, src
    ocfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    refs = [oracle.Circuit(b.netlists[i]).tran(h, N * h, method=method, cfg=ocfg) for i in range(4)]
    x0 = np.stack([xr[0] for xr, _ in refs])
    x, st, it = tran_host_run(src, b.params, c.elements, x0, h, N, method, (1e-3, 1e-12, 1e-9, 300))
    assert (st == 0).all()
    for i, (xr, itr) in enumerate(refs):
        assert np.array_equal(it[i][1:], itr[1:]), i
        assert max(nerr(x[i][k], xr[k]) for k in range(1, N + 1)) < 1e-9


@pytest.mark.parametrize("method", [0, 2])
def test_tran_lane_source_linear_on_host_vs_oracle(oracle, method):
    """Linear circuits through the same kernel: the active filter of every
    controlled source and an RLC ladder (inductor companions)."""
    import krets_amd as kg
    from conftest import active_netlist
    from krets_amd import workloads as W
    for text, h, N in ((active_netlist(), 2e-6, 100),
                       (W.c3_netlist([20.0, 35.0, 50.0], [2e-9, 3e-9, 4e-9], [0.3e-12, 0.5e-12, 0.7e-12]), 1e-11, 200)):
        c = kg.Circuit(text)
        r, src = tran_source(c, c.params, h, method)
        assert r > 0, src
        xr, _ = oracle.Circuit(text).tran(h, N * h, method=method)
        x, st, it = tran_host_run(src, c.params[None], c.elements, xr[:1], h, N, method, (1e-3, 1e-12, 1e-6, 300))
        assert st[0] == 0 and (it[0][1:] == 1).all()
        assert max(nerr(x[0][k], xr[k]) for k in range(1, N + 1)) < 1e-12


def test_tran_lane_static_pivot_guard_on_host(oracle):
    """The generated fixed-step kernel's static-pivot guard on three RLC
    ladders of one topology whose pivots come from the first (L and C of the
    others at 1e-2 and 1e-6 of it): the published error figure of the two
    kept instances equals guard_bound_coeff(m) * || |L||U| ||_inf / ||A||_inf
    of their static-pivot LU, recomputed by dense elimination in numpy (to
    1e-9 relative), their states equal the oracle's, and the third (growth
    above 1e7) reports Decomposition."""
    import krets_amd as kg
    import plan_emulator as pe
    from test_plan import values
    from test_gpu_static_guard import H, netlists
    nets = netlists()
    cs = [kg.Circuit(t) for t in nets]
    N = 40
    for method, alpha in ((1, 1.0), (2, 2.0)):
        r, src = tran_source(cs[0], cs[0].params, H, method)
        assert r > 0, src
        x, st, it, err = tran_host_run(src, np.stack([c.params for c in cs]), cs[0].elements, np.zeros((3, cs[0].n)),
                                       H, N, method, (1e-3, 1e-12, 1e-6, 300), with_err=True)
        assert list(st) == [0, 0, 2], st
        plan = pe.Plan(cs[0].elements, cs[0].n, pe.TRAN, pe.MINDEG, cs[0].params, s_rep=alpha / H)
        for k in (0, 1):
            ref = pe.static_pivot_bound(plan, values(cs[k], pe.TRAN, h=H, alpha=alpha)[0])
            assert abs(err[k] / ref - 1.0) < 1e-9, (method, k, err[k], ref)
            xr, _ = oracle.Circuit(nets[k]).tran(H, N * H, method=method)
            assert max(nerr(x[k][j], xr[j]) for j in range(1, N + 1)) < 1e-9


def test_tran_lane_source_builds_for_gfx950():
    """The fixed-step kernel builds for gfx950 without scratch on the rectifier
    and on the active filter of every controlled source."""
    import krets_amd as kg
    from conftest import active_netlist
    from krets_amd import workloads as W
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    b = W.c4_rectifier(1, seed=4)
    for c in (kg.Circuit(b.netlists[0]), kg.Circuit(active_netlist())):
        r, src = tran_source(c, c.params, 1e-6, 2)
        assert r > 0, src
        d = tempfile.mkdtemp()
        try:
            p = os.path.join(d, "k.hip")
            open(p, "w").write("#include <hip/hip_runtime.h>\n" + src)
            out = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                  "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", p, "-o", p + ".o",
                                  "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True)
            assert out.returncode == 0, out.stderr[-2000:]
            assert int(re.search(r"ScratchSize \[bytes/lane\]: (\d+)", out.stderr).group(1)) == 0
        finally:
            shutil.rmtree(d)


def c1(z):
    """|re| + |im|, the guard's entry norm."""
    return np.abs(np.real(z)) + np.abs(np.imag(z))


def static_growth(plan, c, w):
    """rho = || |L||U| ||_inf / ||A||_inf (|re| + |im| entries) of the plan's
    static-order LU of c's AC matrix at angular frequency w, in numpy."""
    et, val = c.elements["type"], c.params[:, 0]

    def contrib(e, k):
        y = 1.0 / val[e] if et[e] == 0 else 1j * w * val[e]
        return {0: y, 1: -y, 2: 1.0, 3: -1.0, 4: -1j * w * val[e] if et[e] == 2 else -val[e] if et[e] == 0 else 1.0,
                5: val[e], 6: -val[e]}[k]

    _, M = plan.gather(contrib)
    n = M.shape[0]
    L, U = np.eye(n, dtype=complex), M.copy()
    for j in range(n):
        L[j + 1:, j] = U[j + 1:, j] / U[j, j]
        U[j + 1:, :] -= np.outer(L[j + 1:, j], U[j, :])
        U[j + 1:, j] = 0.0
    return np.max(c1(L) @ c1(U).sum(1)) / np.max(c1(M).sum(1))


def guard_bound_coeff(m):
    """The lane kernels' backward-error bound per unit growth (kg_internal.h)."""
    u = 2.0 ** -53
    g2 = 2 * u / (1 - 2 * u)
    q = m * np.sqrt(2) * g2
    return 3.01 * np.sqrt(2) * q / (1 - q)


def test_lane_guard_queues_exactly_the_unstable_solves(oracle):
    """FP64 lane kernel, static pivots chosen at the sweep's centre, on the
    wide RLC batch swept from 1 mHz to 1 GHz: the kernel bounds each solve's
    growth rho = || |L||U| || / ||A|| during its factorisation (from above,
    within a factor 2, through |l_ik| <= |v_ik| |1/u_kk|) and queues it when
    rho > 4. Checked against rho of the same static-order LU in numpy: every
    solve with numpy rho > 4 is queued, every queued one has rho > 2. Every
    kept solve has a normwise backward error below the instance's published
    bound (at most 4 c_m, c_m = 3.01 sqrt(2) g_m) and the oracle's solution to
    1e-12."""
    import krets_amd as kg
    import plan_emulator as pe
    from test_gpu_guard import FREQS, wide_rlc_batch
    nets = wide_rlc_batch()
    cs = [kg.Circuit(t) for t in nets]
    s_rep = 2 * np.pi * np.sqrt(FREQS[0] * FREQS[-1])
    r, src = source(cs[0], 0, s_rep=s_rep)
    assert r > 0, src
    queued = []
    x, st, err = host_run(src, np.stack([c.params for c in cs]), cs[0].elements, FREQS, cs[0].n, queued=queued)
    assert (st == 0).all()
    plan = pe.Plan(cs[0].elements, cs[0].n, pe.AC, pe.MINDEG, cs[0].params, s_rep)
    m = int(np.max(np.diff(plan.row_ptr)))
    nf, kept = len(FREQS), 0
    for i, t in enumerate(nets):
        oc = oracle.Circuit(t)
        ref = oc.ac(FREQS)
        for k, f in enumerate(FREQS):
            rho = static_growth(plan, cs[i], 2 * np.pi * f)
            if i * nf + k in queued:
                assert rho > 2 * (1 - 1e-9), (i, f, rho)
                continue
            kept += 1
            assert rho <= 4 * (1 + 1e-9), (i, f, rho)
            A, b = oc.assemble_ac(np.zeros(oc.n), f)
            xg = x[i, k]
            eta = np.max(np.abs(b - A @ xg)) / (np.max(np.abs(A).sum(1)) * np.max(np.abs(xg)) + np.max(np.abs(b)))
            assert eta <= err[i], (i, f, eta, err[i])
            assert nerr(xg, ref[k]) <= 1e-12, (i, f, nerr(xg, ref[k]))
    assert kept >= 200 and len(queued) >= 16  # the workload does exercise the guard both ways
    assert (err > 0).all() and (err <= 4 * guard_bound_coeff(m) * (1 + 1e-9)).all()


@pytest.mark.parametrize("refine", [2, -2])
def test_lane_mixed_guard_keeps_only_backward_stable_solves(oracle, refine):
    """Mixed precision lane kernels (complex64 factors, FP64 refinement; -2:
    the split-solution form) on the wide RLC batch. Their systems run in
    blocks of 64 instances, frequency-major inside a block, so the queue must
    carry each system as instance * frequencies + frequency index, the id the
    fixup decodes. A solve is queued when its complex64 factorisation grew
    (rho > 4) or its refinement estimate stays above 1e-10: refinement through
    grown factors settles with small corrections at backward errors near
    1e-11. Every kept solve is backward stable: normwise backward error at
    most 2^-46 with an FP64 solution, 5e-14 in the split form (its solution
    is a complex64 first solution plus a complex64 sum of corrections)."""
    import krets_amd as kg
    import plan_emulator as pe
    from test_gpu_guard import FREQS, wide_rlc_batch
    nets = wide_rlc_batch()
    cs = [kg.Circuit(t) for t in nets]
    s_rep = 2 * np.pi * np.sqrt(FREQS[0] * FREQS[-1])
    r, src = source(cs[0], 1, refine=refine, s_rep=s_rep)
    assert r > 0, src
    queued = []
    x, st, err = host_run(src, np.stack([c.params for c in cs]), cs[0].elements, FREQS, cs[0].n,
                          refine=abs(refine), queued=queued)
    assert (st == 0).all()
    plan = pe.Plan(cs[0].elements, cs[0].n, pe.AC, pe.MINDEG, cs[0].params, s_rep)
    nf = len(FREQS)
    assert len(set(queued)) == len(queued) and all(0 <= q < len(nets) * nf for q in queued)
    bound = 2.0 ** -46 if refine > 0 else 5e-14
    grown = 0
    for i, t in enumerate(nets):
        oc = oracle.Circuit(t)
        for k, f in enumerate(FREQS):
            rho = static_growth(plan, cs[i], 2 * np.pi * f)
            if i * nf + k in queued:
                continue
            grown += rho > 2
            A, b = oc.assemble_ac(np.zeros(oc.n), f)
            xg = x[i, k]
            eta = np.max(c1(b - A @ xg)) / (np.max(c1(A).sum(1)) * np.max(c1(xg)) + np.max(c1(b)))
            assert eta <= bound, (i, f, eta)
            assert rho <= 4 * (1 + 1e-6), (i, f, rho)  # complex64 and FP64 factors grow alike
    assert len(queued) >= 16 and grown > 0


HIPRTC = "/opt/rocm/lib/libhiprtc.so"


def hiprtc_compile(src):
    """Compiles a generated source with hiprtc and the library's own options
    (jit.cpp: jit_load) for gfx950; no GPU is needed. Returns (ok, log)."""
    L = ctypes.CDLL(HIPRTC)
    prog = ctypes.c_void_p()
    assert L.hiprtcCreateProgram(ctypes.byref(prog), src.encode(), b"kg_jit.hip", 0, None, None) == 0
    try:
        opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-Xclang", b"-target-feature", b"-Xclang",
                b"-packed-fp32-ops"]
        rc = L.hiprtcCompileProgram(prog, len(opts), (ctypes.c_char_p * len(opts))(*opts))
        n = ctypes.c_size_t()
        L.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
        log = ctypes.create_string_buffer(n.value + 1)
        L.hiprtcGetProgramLog(prog, log)
        return rc == 0, log.value.decode(errors="replace")
    finally:
        L.hiprtcDestroyProgram(ctypes.byref(prog))


@pytest.mark.skipif(not os.path.exists(HIPRTC), reason="hiprtc not installed")
def test_every_generated_form_compiles_with_hiprtc():
    """The library compiles its specialised kernels with hiprtc at run time,
    which has neither the HIP runtime headers nor the C library's macros that
    hipcc builds see: every generated form goes through the same compiler
    here, on the CPU, before any GPU run."""
    import krets_amd as kg
    from krets_amd import workloads as W
    lad = kg.Circuit(W.c2_rc_ladder_mc(1, sections=12, seed=2).netlists[0])
    act = kg.Circuit(active_netlist())
    rect = kg.Circuit(W.c4_rectifier(1, seed=4).netlists[0])
    forms = []
    for c in (lad, act):
        for precision, refine in ((0, 2), (1, 2), (1, -1)):
            forms.append((f"ac {precision} {refine}", source(c, precision, refine)))
        forms.append(("dc", dc_source(c, c.params)))
    for c in (lad, rect):
        forms.append(("lte", lte_source(c, c.params, 1e-9)))
        for method in (0, 1, 2):
            forms.append((f"tran {method}", tran_source(c, c.params, 1e-9, method)))
    built = 0
    for name, (r, src) in forms:
        if r <= 0:  # a form the generator declines for this circuit
            continue
        ok, log = hiprtc_compile(src)
        assert ok, (name, log[-2000:])
        built += 1
    assert built >= 10
