"""GPU parity: libkrets_gpu.so (HIP, gfx950) against the CPU oracle.

Every case runs the HIP kernels through the C ABI and compares with the
oracle (krets semantics) on the same inputs. Bounds are normwise relative
per solve, ||x - x_ref||_2 / ||x_ref||_2: 1e-12 for the series+shunt ladder
and linear OP cases, 1e-9 elsewhere (mixed precision included); the krets
known-answer values are asserted at krets' own tolerances as well.
"""
import math
import sys

import numpy as np
import pytest

from conftest import read_circuit

pytestmark = pytest.mark.gpu

TPS = [1, 4, 16, 64, 256]


def say(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def nerr(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    d = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (d if d > 0 else 1.0)


@pytest.fixture(scope="module")
def kg():
    import krets_amd
    return krets_amd


def gpu_op(kg, text, tps=0, compat=True):
    c = kg.Circuit(text)
    s = kg.Solver(c, kg.SolverConfig(threads_per_system=tps, compat=compat))
    return c, s.solve(kg.Op())


OP_CASES = [
    "circuit_simulation_farid_n_najm/circuit_simulation_farid_n_najm.cir",
    "voltage_divider/voltage_divider.cir",
    "low_pass_filter/low_pass_filter.cir",
    "high_pass_filter/high_pass_filter.cir",
    "basic_001/basic_001.cir",
    "resistor_ladder_500/resistor_ladder_500.cir",
    "diode_iv_curve/diode_iv_curve.cir",
]


@pytest.mark.parametrize("tps", TPS)
@pytest.mark.parametrize("path", OP_CASES)
def test_op_vs_oracle(kg, oracle, path, tps):
    say(f"op {path} tps={tps}")
    text = read_circuit(path)
    c, r = gpu_op(kg, text, tps)
    oc = oracle.Circuit(text)
    xr, _ = oc.op()
    x = np.array([r[k] for k in oc.names])
    tol = 1e-9 if "diode" in path else 1e-12
    assert nerr(x, xr) < tol


def test_op_larger_than_lds(kg, oracle):
    """krets' resistor_ladder_5000 (a series chain of 5000 resistors from V1 to
    ground, benches/benchmark.rs:38-51): 5001 unknowns need 480 KB per system,
    more than a CU's LDS, so the operating point runs with its values in a
    global workspace. V(k) = 1 - (k - 1) / 5000 and I(V1) = -1 / 5 MOhm.
    The chain's conductance matrix (a path Laplacian) has condition number
    about 4 n^2 / pi^2 = 1e7, so rounding alone allows errors near
    kappa u = 1e-9 of the 1 V scale; the bound is that (measured: 3.5e-11 on
    the GPU, 1.6e-12 for the oracle's different elimination order)."""
    from krets_amd import workloads as W
    say("OP larger than LDS (5000-resistor chain)")
    text = W.series_ladder(5000)
    c, r = gpu_op(kg, text, 0)
    v = np.array([r[f"V({k})"] for k in range(1, 5001)])
    assert np.max(np.abs(v - (1 - np.arange(5000) / 5000.0))) < 1e-9
    assert abs(r["I(V1)"] / (-1.0 / 5e6) - 1) < 1e-9
    xr, _ = oracle.Circuit(text).op()
    assert nerr(np.array([r[k] for k in c.names]), xr) < 1e-9


@pytest.mark.parametrize("path", [p for p in OP_CASES if "diode" not in p and "ladder" not in p])
def test_op_lane_kernel_vs_oracle(kg, oracle, path):
    """Linear operating points through the run-time specialised lane kernel
    (2001), which the default mapping also takes for these circuits; a
    system too large for one lane's registers is refused when 2001 is asked
    for and solved by the level-scheduled kernel by default."""
    say(f"op lane {path}")
    text = read_circuit(path)
    c, r = gpu_op(kg, text, 2001)
    xr, _ = oracle.Circuit(text).op()
    assert nerr(np.array([r[k] for k in oracle.Circuit(text).names]), xr) < 1e-12
    if "divider" in path:
        big = read_circuit("resistor_ladder_500/resistor_ladder_500.cir")
        with pytest.raises(kg.KretsGpuError) as e:
            gpu_op(kg, big, 2001)
        assert e.value.code == 6
        _, r = gpu_op(kg, big, 0)
        assert abs(r["I(V1)"] + 1 / (500 * 1000.0)) < 1e-15


def test_op_krets_known_answers(kg):
    say("op known answers")
    _, s = gpu_op(kg, read_circuit("circuit_simulation_farid_n_najm/circuit_simulation_farid_n_najm.cir"))
    exp = {"V(4)": 1.9888, "V(8)": 1.0, "V(3)": 2.00879, "V(2)": 1.80879, "V(6)": 1.98814, "V(5)": 2.0,
           "V(1)": 1.88527, "V(7)": 3.98814, "I(V3)": 0.0, "I(V2)": -199.88e-3, "I(V1)": -198.88e-3}
    for k, v in exp.items():  # test_op.rs:25-46
        assert abs(s[k] - v) < 1e-3
    _, s = gpu_op(kg, "V1 2 1 32\nR1 1 0 2\nR2 2 3 4\nR3 2 0 8\nV2 3 0 20\n")  # test_op.rs:48-68
    for k, v in {"V(1)": -8.0, "V(2)": 24.0, "V(3)": 20.0, "I(V1)": -4.0, "I(V2)": 1.0}.items():
        assert abs(s[k] - v) < 1e-12
    _, s = gpu_op(kg, read_circuit("diode_iv_curve/diode_iv_curve.cir"))  # test_op.rs:131-140
    assert abs(s["V(out)"] - 0.517) < 1e-3


@pytest.mark.parametrize("tps", [0, 1, 64, 256, 512, 1024])  # 0: the default, 512 threads for this single system
def test_c1_series_shunt_ladder(kg, oracle, tps):
    from krets_amd import workloads as W
    say(f"C1 ladder tps={tps}")
    text = W.c1_ladder()
    c, r = gpu_op(kg, text, tps)
    oc = oracle.Circuit(text)
    xr, _ = oc.op()
    x = np.array([r[k] for k in oc.names])
    assert nerr(x, xr) < 1e-12
    phi = (1 + math.sqrt(5)) / 2  # I(V1) = -phi mA (infinite-ladder closed form)
    assert abs(r["I(V1)"] / (-phi * 1e-3) - 1) < 1e-12
    text = read_circuit("resistor_ladder_1000/resistor_ladder_1000.cir")
    c, r = gpu_op(kg, text, tps)
    v = np.array([r[f"V({k})"] for k in range(1, 1001)])
    assert np.max(np.abs(v - (1 - np.arange(1000) / 1000.0))) < 1e-12


def test_dc_sweep(kg, oracle):
    say("dc sweep divider + diode")
    text = read_circuit("voltage_divider/voltage_divider.cir")
    s = kg.Solver(kg.Circuit(text))
    res = s.solve(kg.Dc("V1", 0.0, 1.0, 0.1))
    xr = oracle.Circuit(text).dc("V1", 0.0, 1.0, 0.1)
    names = oracle.Circuit(text).names
    assert len(res) == xr.shape[0] == 11
    for p, rr in enumerate(res):
        assert nerr([rr[k] for k in names], xr[p]) < 1e-12
        assert rr["step"] == float(p)
    assert abs(res[-1]["I(V1)"] + 1 / 3000) < 1e-4  # test_dc.rs:53
    text = read_circuit("diode_iv_curve/diode_iv_curve.cir")
    oc = oracle.Circuit(text)
    xr = oc.dc("V1", 0.0, 2.0, 0.25)
    res = kg.Solver(kg.Circuit(text)).solve(kg.Dc("V1", 0.0, 2.0, 0.25))
    for p, rr in enumerate(res):
        assert nerr([rr[k] for k in oc.names], xr[p]) < 1e-9


@pytest.mark.parametrize("precision", [0, 1])
def test_ac_known_answers_and_golden(kg, oracle, precision):
    import json, os
    from conftest import GOLDEN
    say(f"ac golden precision={precision}")
    s = kg.Solver(kg.Circuit(read_circuit("low_pass_filter/low_pass_filter.cir")), precision=precision)
    r = s.solve(kg.Ac(kg.Linear(1), 1000.0, 1000.0))[0]
    assert abs(r["V(out)"] - (2.470452e-02 - 1.55223e-01j)) < 1e-6  # test_ac.rs:25-46
    s = kg.Solver(kg.Circuit(read_circuit("high_pass_filter/high_pass_filter.cir")), precision=precision)
    r = s.solve(kg.Ac(kg.Linear(1), 1000.0, 1000.0))[0]
    assert r["frequency"] == 1000.0
    assert abs(r["V(out)"] - (7.169568e-01 - 4.50477e-01j)) < 1e-6  # test_ac.rs:48-74
    assert abs(r["I(V1)"] - (-7.16957e-03 + 4.504772e-03j)) < 1e-8
    g = json.load(open(os.path.join(GOLDEN, "low_pass_ac_result.json")))
    cols = {k: np.array([float(v) for v in vals]) for k, vals in g["columns"].items()}
    s = kg.Solver(kg.Circuit(read_circuit("low_pass_filter/ac/ac.cir")), precision=precision)
    res = s.solve(kg.Ac(kg.Linear(10), 0.1, 1000.0))
    assert np.array_equal([r["frequency"].real for r in res], cols["frequency"])
    tol = 1e-12 if precision == 0 else 1e-9
    for name in ("I(V1)", "V(in)", "V(out)"):
        v = np.array([r[name] for r in res])
        assert np.max(np.abs(np.abs(v) - cols[name + "_mag"]) / cols[name + "_mag"]) < tol
        ph = np.degrees(np.arctan2(v.imag, v.real))
        assert np.max(np.abs(ph - cols[name + "_phase_deg"])) < 1e3 * tol


TRAN_CASES = [("low_pass_filter/transient.cir", 50e-6, 20e-3), ("high_pass_filter/transient.cir", 10e-6, 2e-3),
              ("dual_rc_ladder/dual_rc_ladder.cir", 50e-6, 50e-3), ("rectifier/rectifier.cir", 50e-6, 50e-3)]


@pytest.mark.parametrize("path,h,stop", TRAN_CASES)
def test_transient_compat_vs_oracle(kg, oracle, path, h, stop):
    say(f"transient compat {path}")
    text = read_circuit(path)
    oc = oracle.Circuit(text)
    xr, itr = oc.tran(h, stop)
    res = kg.Solver(kg.Circuit(text)).solve(kg.Transient(h, stop))
    x = np.array([[r[k] for k in oc.names] for r in res])
    assert x.shape == xr.shape
    errs = [nerr(x[i], xr[i]) for i in range(len(x))]
    assert max(errs) < 1e-9, max(errs)
    if "low_pass" in path:  # test_transient.rs:203-206 (the source-lag pin)
        assert abs(res[42]["V(out)"] - 0.8647) < 1e-3 and abs(res[42]["time"] - 2.1e-3) < 1e-6


# 1000 + t: persistent LDS-staged variant; 2001: lane per system, run-time
# specialised; 2002: as 2001 with the solution split into x0 (LDS) + dx;
# 2011 / 2012: both on the balanced elimination order
@pytest.mark.parametrize("tps", TPS + [1004, 1016, 1064, 2001, 2002, 2011, 2012])
@pytest.mark.parametrize("precision", [0, 1])
def test_batched_ac_rc_ladder(kg, oracle, precision, tps):
    from krets_amd import workloads as W
    say(f"batched AC precision={precision} tps={tps}")
    b = W.c2_rc_ladder_mc(24, sections=50, seed=11)
    f = W.c2_frequencies(200)[::10]
    ctx = kg.Context()
    b.load(ctx)
    if tps in (2002, 2012) and precision == 0:
        with pytest.raises(kg.KretsGpuError):
            ctx.ac_sweep(f, precision=precision, config=kg.SolverConfig(threads_per_system=tps))
        return
    xs, st, err = ctx.ac_sweep(f, precision=precision, config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all()
    for i in (0, 7, 23):
        oc = oracle.Circuit(b.netlists[i])
        xr = oc.ac(f)
        e = max(nerr(xs[i][j], xr[j]) for j in range(len(f)))
        assert e < (1e-12 if precision == 0 else 1e-9), (i, e)
    if precision == 1:
        assert (err < 1e-8).all()  # a-posteriori estimate of the remaining refinement error


@pytest.mark.parametrize("lane", [2001, 2002])
def test_lane_ac_mixed_across_instance_blocks(kg, oracle, lane):
    """The mixed lane kernels read element values from the table blocked by 64
    instances: 130 instances are two full blocks and a partial one of 2, with
    an odd frequency count. Every result is checked against the FP64 lane
    kernel, and instances at the block edges against the oracle."""
    from krets_amd import workloads as W
    say(f"mixed lane AC across instance blocks tps={lane}")
    b = W.c2_rc_ladder_mc(130, sections=50, seed=29)
    f = W.c2_frequencies(200)[::12]
    ctx = kg.Context()
    b.load(ctx)
    x64, st64, _ = ctx.ac_sweep(f, precision=0, config=kg.SolverConfig(threads_per_system=2001))
    xm, st, err = ctx.ac_sweep(f, precision=1, config=kg.SolverConfig(threads_per_system=lane))
    assert (st == 0).all() and (st64 == 0).all()
    e = np.linalg.norm(xm - x64, axis=2) / np.linalg.norm(x64, axis=2)
    assert e.max() < 1e-9, e.max()
    assert (err <= 1e-9).all()
    for i in (0, 63, 64, 127, 128, 129):
        xr = oracle.Circuit(b.netlists[i]).ac(f)
        assert max(nerr(xm[i][j], xr[j]) for j in range(len(f))) < 1e-9, i


@pytest.mark.parametrize("refine", [1, 2, 3])
@pytest.mark.parametrize("precision,lane", [(0, 2001), (1, 2001), (1, 2002), (0, 2011), (1, 2012)])
def test_lane_ac_golden_and_rlc(kg, oracle, precision, lane, refine):
    """Run-time specialised lane kernel: the krets AC golden vector and a shared
    RLC ladder batch (inductor branch rows) vs the oracle."""
    import json, os
    from conftest import GOLDEN
    from krets_amd import workloads as W
    say(f"lane AC precision={precision} lane={lane} refine={refine}")
    cfg = kg.SolverConfig(threads_per_system=lane, refine_iters=refine)
    g = json.load(open(os.path.join(GOLDEN, "low_pass_ac_result.json")))
    cols = {k: np.array([float(v) for v in vals]) for k, vals in g["columns"].items()}
    c = kg.Circuit(read_circuit("low_pass_filter/ac/ac.cir"))
    ctx = kg.Context()
    ctx.load(c.elements, c.params[None], c.n)
    f = np.array(cols["frequency"])
    xs, st, _ = ctx.ac_sweep(f, precision=precision, config=cfg)
    assert (st == 0).all()
    tol = 1e-12 if precision == 0 else 1e-9
    for name in ("I(V1)", "V(in)", "V(out)"):
        v = xs[0][:, c.names.index(name)]
        assert np.max(np.abs(np.abs(v) - cols[name + "_mag"]) / cols[name + "_mag"]) < tol
    rng = np.random.default_rng(5)
    nets = []
    for _ in range(6):
        lines = ["V1 n0 0 0 AC 1"]
        for k in range(1, 9):
            lines += [f"R{k} n{k-1} m{k} {rng.uniform(10, 100)!r}", f"L{k} m{k} n{k} {rng.uniform(1e-9, 1e-8)!r}",
                      f"C{k} n{k} 0 {rng.uniform(1e-13, 1e-12)!r}"]
        nets.append("\n".join(lines) + "\n")
    cs = [kg.Circuit(t) for t in nets]
    ctx.load(cs[0].elements, np.stack([x.params for x in cs]), cs[0].n)
    f = np.geomspace(1e6, 1e10, 37)
    xs, st, err = ctx.ac_sweep(f, precision=precision, config=cfg)
    assert (st == 0).all()
    for i in range(6):
        xr = oracle.Circuit(nets[i]).ac(f)
        e = max(nerr(xs[i][j], xr[j]) for j in range(len(f)))
        assert e < (1e-12 if precision == 0 else (1e-9 if refine >= 2 else 1e-6)), (i, e)


@pytest.mark.parametrize("tps", [1, 16, 64, 256, 1064, 1256])  # 1000 + t: plan staged in LDS
@pytest.mark.parametrize("method,precision", [(2, 0), (2, 1), (1, 0), (0, 0)])
def test_batched_transient_uneven_rlc(kg, oracle, method, precision, tps):
    from krets_amd import workloads as W
    say(f"batched transient method={method} precision={precision} tps={tps}")
    b = W.c3_rlc_uneven(5, smin=8, smax=40, seed=13)
    ctx = kg.Context()
    b.load(ctx)
    h, n = 10e-12, 300
    xs, st, (it, err) = ctx.transient(h, n, method=method, precision=precision,
                                      config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all()
    for i in range(5):
        oc = oracle.Circuit(b.netlists[i])
        xr, _ = oc.tran(h, n * h, method=method)
        e = max(nerr(xs[i][k], xr[k]) for k in range(1, n + 1))
        assert e < 1e-9, (i, e)


@pytest.mark.parametrize("tps", [3064, 3128, 3256, 3512, 4024])  # 3000 + t: whole step loop in LDS
@pytest.mark.parametrize("method", [2, 1, 0])
def test_lds_resident_transient_uneven_rlc(kg, oracle, method, tps):
    from krets_amd import workloads as W
    say(f"LDS-resident transient method={method} tps={tps}")
    b = W.c3_rlc_uneven(5, smin=8, smax=40, seed=13)
    ctx = kg.Context()
    b.load(ctx)
    h, n = 10e-12, 300
    xs, st, (it, err) = ctx.transient(h, n, method=method, config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all() and (it[:, 1:] == 1).all() and (it[:, 0] == 0).all()
    assert (err > 0).all() and (err < 1e-13).all(), err  # the static-pivot bound of these ladders
    for i in range(5):
        xr, _ = oracle.Circuit(b.netlists[i]).tran(h, n * h, method=method)
        e = max(nerr(xs[i][k], xr[k]) for k in range(n + 1))
        assert e < 1e-12, (i, e)
    with pytest.raises(kg.KretsGpuError):  # FP64 only
        ctx.transient(h, n, method=method, precision=kg.MIXED, config=kg.SolverConfig(threads_per_system=tps))


def test_lds_resident_transient_largest_ladder(kg, oracle):
    """512 RLC sections (1538 unknowns): the compact plan, factors and state
    use 158.7 KB of the CU's 160 KB of LDS."""
    say("LDS-resident transient, 512-section ladder")
    from krets_amd import workloads as W
    b = W.c3_rlc_uneven(2, smin=512, smax=512, seed=19)
    ctx = kg.Context()
    b.load(ctx)
    h, n = 10e-12, 60
    xs, st, _ = ctx.transient(h, n, method=kg.TRAP, config=kg.SolverConfig(threads_per_system=3512))
    assert (st == 0).all()
    for i in range(2):
        xr, _ = oracle.Circuit(b.netlists[i]).tran(h, n * h, method=oracle.TRAP)
        assert max(nerr(xs[i][k], xr[k]) for k in range(n + 1)) < 1e-12


@pytest.mark.parametrize("precision", [0, 1])
def test_bucketed_transient_uneven_rlc(kg, oracle, precision):
    """Autotuned uneven batch of 96 ladders, 8..300 sections: instances run in
    size-class launches on concurrent streams; a sample of each class vs the
    oracle, and the whole batch vs the single-launch mapping."""
    from krets_amd import workloads as W
    say(f"bucketed transient precision={precision}")
    b = W.c3_rlc_uneven(96, smin=8, smax=300, seed=17)
    ctx = kg.Context()
    b.load(ctx)
    h, n = 10e-12, 120
    xs, st, (it, err) = ctx.transient(h, n, method=kg.TRAP, precision=precision)
    assert (st == 0).all()
    ys, st2, _ = ctx.transient(h, n, method=kg.TRAP, precision=precision,
                               config=kg.SolverConfig(threads_per_system=256))
    assert (st2 == 0).all()
    for i in range(96):
        assert max(nerr(xs[i][k], ys[i][k]) for k in range(1, n + 1)) < 1e-12, i
    order = np.argsort(b.n_unknowns)
    for i in (order[0], order[40], order[-1]):
        xr, _ = oracle.Circuit(b.netlists[i]).tran(h, n * h, method=oracle.TRAP)
        e = max(nerr(xs[i][k], xr[k]) for k in range(1, n + 1))
        assert e < 1e-9, (i, e)


@pytest.mark.parametrize("precision", [0, 1])
def test_ac_graph_replay(kg, oracle, precision):
    """Repeated sweeps replay one captured graph (operating point, fills,
    sweep kernel). Every replay, with eager work and synchronisations in
    between, must equal the eager result bit for bit; new parameters and new
    output buffers must be honoured."""
    import torch
    from krets_amd import workloads as W
    say(f"AC graph replay precision={precision}")
    b = W.c2_rc_ladder_mc(40, sections=50, seed=23)
    f = W.c2_frequencies(200)[::4]
    ctx = kg.Context()
    b.load(ctx)
    ctx.use_graphs(False)
    ref, st, err_ref = ctx.ac_sweep(f, precision=precision)
    assert (st == 0).all()
    ctx.use_graphs(True)
    dev = torch.device("cuda:0")
    out = torch.empty(40 * len(f) * b.n_unknowns * 2, dtype=torch.float64, device=dev)
    for k in range(5):
        out.fill_(float("nan"))
        _, st, err = ctx.ac_sweep(f, precision=precision, out=out)
        assert (st == 0).all()
        x = out.view(40, len(f), b.n_unknowns, 2).cpu().numpy()
        for i in (0, 17, 39):
            assert np.array_equal(x[i, ..., 0] + 1j * x[i, ..., 1], ref[i]), (k, i)
        assert np.array_equal(err, err_ref)
        ctx.op()  # eager work on the same stream between replays
        torch.cuda.synchronize()
    # new parameter values: the replayed graph reads them from device memory
    p2 = b.params.copy()
    p2[:, :, 0] *= 1.01
    ctx.load(b.elements, p2, b.n_unknowns)
    x2, st, _ = ctx.ac_sweep(f, precision=precision)
    x3, st3, _ = ctx.ac_sweep(f, precision=precision)
    assert (st == 0).all() and (st3 == 0).all()
    assert np.array_equal(x2[5], x3[5])
    xr = oracle.Circuit(W.c2_netlist(p2[5, 1::2, 0], p2[5, 2::2, 0])).ac(f)  # elements: V1, R1, C1, R2, ...
    e = max(nerr(x3[5][j], xr[j]) for j in range(len(f)))
    assert e < (1e-12 if precision == 0 else 1e-9), e


def test_error_statuses(kg):
    say("error statuses")
    s = kg.Solver(kg.Circuit("I1 1 0 1\nR1 1 0 1\n"))
    with pytest.raises(kg.KretsGpuError) as e:
        s.solve(kg.Ac(kg.Linear(1), 1.0, 1.0))
    assert e.value.code == 2
    s = kg.Solver(kg.Circuit(read_circuit("diode_iv_curve/diode_iv_curve.cir")),
                  kg.SolverConfig(maximum_iterations=1))
    with pytest.raises(kg.KretsGpuError) as e:
        s.solve(kg.Op())
    assert e.value.code == 3
    s = kg.Solver(kg.Circuit("V1 1 0 1\nR1 1 2 1\nC1 2 0 1u G2\n"))
    with pytest.raises(kg.KretsGpuError) as e:
        s.solve(kg.Op())
    assert e.value.code == 2
    s = kg.Solver(kg.Circuit(read_circuit("voltage_divider/voltage_divider.cir")))
    with pytest.raises(kg.KretsGpuError) as e:
        s.solve(kg.Dc("V9", 0, 1, 1))
    assert e.value.code == 4


def test_lane_mapping_errors_and_singular_lanes(kg):
    """The run-time specialised lane mappings (2001) refuse mixed precision and
    per-instance topologies with InvalidArgument, and a lane whose pivot is
    zero reports Decomposition for its own instance only."""
    from krets_amd import workloads as W
    say("lane mapping errors")
    b = W.c4_rectifier(4, seed=4)
    ctx = kg.Context()
    b.load(ctx)
    with pytest.raises(kg.KretsGpuError) as e:
        ctx.transient(1e-6, 10, method=kg.BE, precision=kg.MIXED, config=kg.SolverConfig(threads_per_system=2001))
    assert e.value.code == 9
    t = W.c3_rlc_uneven(3, smin=8, smax=16, seed=7)
    ctx.load(t.elements, t.params, t.n_unknowns, t.elem_offset)
    with pytest.raises(kg.KretsGpuError) as e:
        ctx.transient(1e-11, 10, method=kg.TRAP, config=kg.SolverConfig(threads_per_system=2001))
    assert e.value.code == 9
    # instance 1 has a zero resistor (infinite conductance): only its status fails
    c = kg.Circuit("V1 1 0 1\nR1 1 2 1k\nR2 2 0 1k\n")
    p = np.repeat(c.params[None], 3, axis=0)
    p[1, 1, 0] = 0.0
    ctx.load(c.elements, p, c.n)
    x, st, _ = ctx.op(config=kg.SolverConfig(threads_per_system=0))
    assert list(st) == [0, 2, 0] and abs(x[0][c.names.index("V(2)")] - 0.5) < 1e-15


@pytest.mark.parametrize("tps", [0, 1, 4, 2001])  # 2001: lane kernel specialised at run time
def test_lte_rectifier_vs_oracle(kg, oracle, tps):
    from conftest import time_aligned
    from krets_amd import workloads as W
    say(f"LTE rectifier tps={tps}")
    b = W.c4_rectifier(16, seed=4)
    ctx = kg.Context()
    b.load(ctx)
    cfg = kg.SolverConfig(compat=False, voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12,
                          threads_per_system=tps)
    args = (5e-3, 1e-6, 1e-14, 5e-5, 1e-3, 1e-6)
    t, xs, npts, nt, rej = ctx.transient_lte(*args, 4000, cfg)
    ocfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    for i in (0, 5, 15):
        tr, xr, ntr, rejr = oracle.Circuit(b.netlists[i]).tran_lte(*args, cfg=ocfg)
        # the same step controller: the same accepted grid, Newton and rejection counts
        assert npts[i] == len(tr) and nt[i] == ntr and rej[i] == rejr, (i, npts[i], len(tr), nt[i], ntr)
        assert np.max(np.abs(t[i][:npts[i]] - tr)) <= 1e-9 * 5e-3  # rounding of the accumulated time
        xa = time_aligned(tr, xr, t[i][:npts[i]])
        e = max(nerr(xs[i][j], xa[j]) for j in range(len(tr)))
        assert e < 1e-9, (i, e)


@pytest.mark.parametrize("tps", [0, 1, 64, 2001])  # 2001: lane kernel specialised at run time
def test_fixed_step_compat_rectifier_instances(kg, oracle, tps):
    from krets_amd import workloads as W
    say(f"fixed-step compat rectifier tps={tps}")
    b = W.c4_rectifier(8, seed=4)
    ctx = kg.Context()
    b.load(ctx)
    cfg = kg.SolverConfig(voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12, threads_per_system=tps)
    xs, st, (it, _) = ctx.transient(1e-6, 2000, method=kg.BE_COMPAT, config=cfg)
    assert (st == 0).all()
    for i in (0, 3, 7):
        xr, itr = oracle.Circuit(b.netlists[i]).tran(1e-6, 2e-3, cfg=oracle.config(vabstol=1e-9, iabstol=1e-12))
        assert np.array_equal(it[i], itr)
        # normwise relative error with the norm floored at the voltage tolerance (1e-9 V): the operating
        # point of a 0 V source is exactly 0, and both codes return rounding noise of ~1e-34 there
        errs = [np.linalg.norm(xs[i][k] - xr[k]) / max(np.linalg.norm(xr[k]), 1e-9) for k in range(xr.shape[0])]
        k = int(np.argmax(errs))
        print(f"instance {i}: max normwise {errs[k]:.3e} at step {k}: gpu {xs[i][k]} oracle {xr[k]} "
              f"iters {it[i][max(k-2,0):k+2]}", flush=True)
        assert errs[k] < 1e-9, (i, errs[k])
