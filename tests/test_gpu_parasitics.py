"""GPU: the opt-in diode parasitics (series resistance, junction charge)
against the oracle.

Netlists are read with KG_NETLIST_DIODE_PARASITICS on both sides (the
library's kg_parse_netlist_ext, the oracle's ref_parse_ext). krets itself
ignores rs and has no junction charge (crates/krets-parser/src/models/
diode.rs:14-37), so the oracle's restatement of both is what the kernels are
held to; it is pinned against closed forms and an ODE solve in
tests/test_parasitics.py. Bounds: normwise relative per solve 1e-9, with the
norm floored at 1e-9 where a solution can be exactly zero; the same accepted
step grid and Newton and rejection counts for the step-size controller.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VT = 1.380649e-23 * 300.0 / 1.602176634e-19


def say(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def nerr(x, ref):
    return np.linalg.norm(np.asarray(x) - ref) / max(np.linalg.norm(ref), 1e-9)


def model(rng, rs=True):
    cjo = float(rng.uniform(0.5, 5)) * 1e-12
    tt = float(rng.uniform(0, 10)) * 1e-9
    r = f" rs={float(rng.uniform(1, 50))!r}" if rs else ""
    return f".model dm D(is=1e-14 n={float(rng.uniform(1, 2))!r}{r} cjo={cjo!r} vj=0.7 m={float(rng.uniform(0.3, 0.5))!r} fc=0.5 tt={tt!r})\n"


def rectifier_batch(ni, seed, amp=2.0, f=20e6):
    """SIN-driven series R into a diode (with rs) loaded by R || C."""
    rng = np.random.default_rng(seed)
    nets = []
    for _ in range(ni):
        nets.append(f"V1 in 0 SIN(0 {amp} {f} 0 0 0)\nR1 in a {float(rng.uniform(50, 200))!r}\nD1 a out dm\n"
                    f"RL out 0 {float(rng.uniform(1e3, 1e4))!r}\nCL out 0 {float(rng.uniform(1, 10)) * 1e-12!r}\n"
                    + model(rng))
    return nets


def load(kg, nets):
    cs = [kg.Circuit(t, diode_parasitics=True) for t in nets]
    ctx = kg.Context()
    ctx.load(cs[0].elements, np.stack([c.params for c in cs]), cs[0].n)
    return cs, ctx


@pytest.mark.parametrize("tps", [0, 1, 16, 64])
def test_op_with_series_resistance(oracle, tps):
    import krets_amd as kg
    say(f"parasitics OP tps={tps}")
    rng = np.random.default_rng(7)
    nets = [f"V1 in 0 {float(v)!r}\nR1 in a 1k\nD1 a 0 dm\nD2 a 0 dm\n" + model(rng) for v in rng.uniform(0.5, 5, 8)]
    cs, ctx = load(kg, nets)
    x, st, _ = ctx.op(config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all(), st
    for i, t in enumerate(nets):
        xr, _ = oracle.Circuit(t, diode_parasitics=True).op()
        assert nerr(x[i], xr) < 1e-9, (i, nerr(x[i], xr))


@pytest.mark.parametrize("precision,tps", [(0, 0), (0, 16), (0, 64), (0, 1016), (1, 0), (1, 16)])
def test_ac_with_junction_capacitance(oracle, precision, tps):
    """Biased diodes (reverse and forward, both sides of fc vj) behind R: the
    small-signal junction capacitance at the operating point."""
    import krets_amd as kg
    say(f"parasitics AC precision={precision} tps={tps}")
    rng = np.random.default_rng(11)
    nets = [f"V1 in 0 DC {float(v)!r} AC 1\nR1 in a 1k\nD1 a 0 dm\nC1 a 0 1p\n" + model(rng)
            for v in np.linspace(-3, 0.8, 12)]
    cs, ctx = load(kg, nets)
    f = np.logspace(3, 9.5, 40)
    xs, st, err = ctx.ac_sweep(f, precision=precision, config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all(), st
    worst = 0.0
    for i, t in enumerate(nets):
        xr = oracle.Circuit(t, diode_parasitics=True).ac(f)
        worst = max(worst, max(nerr(xs[i][k], xr[k]) for k in range(len(f))))
    assert worst < 1e-9, worst


@pytest.mark.parametrize("method", ["BE", "TRAP"])
@pytest.mark.parametrize("tps", [0, 1, 16, 64])
def test_transient_with_junction_charge(oracle, method, tps):
    import krets_amd as kg
    say(f"parasitics transient {method} tps={tps}")
    nets = rectifier_batch(16, seed=5)
    cs, ctx = load(kg, nets)
    h, n = 0.05e-9, 1000
    m = getattr(kg, method)
    cfg = kg.SolverConfig(compat=False, voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12,
                          threads_per_system=tps)
    xs, st, (it, _) = ctx.transient(h, n, method=m, config=cfg)
    assert (st == 0).all(), st
    ocfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    for i in (0, 7, 15):
        xr, itr = oracle.Circuit(nets[i], diode_parasitics=True).tran(h, h * n, method=getattr(oracle, method),
                                                                      cfg=ocfg)
        assert np.array_equal(it[i], itr), i
        e = max(nerr(xs[i][k], xr[k]) for k in range(xr.shape[0]))
        assert e < 1e-9, (i, e)


def test_transient_lane_kernel_declines_junction_charge():
    import krets_amd as kg
    say("parasitics lane kernel declines")
    cs, ctx = load(kg, rectifier_batch(4, seed=5))
    with pytest.raises(kg.KretsGpuError):
        ctx.transient(0.05e-9, 10, method=kg.TRAP, config=kg.SolverConfig(compat=False, threads_per_system=2001))


@pytest.mark.parametrize("tps", [0, 1])
def test_lte_with_junction_charge(oracle, tps):
    """The step-size controller runs the oracle's algorithm: the same numbers
    of accepted points, Newton iterations and rejections, on grids that agree
    to 1e-6 of the interval. Its decisions amplify rounding: on instance 7 a
    1e-15 relative change of one resistor moves accepted times by 6e-16 s, and
    an interpolated comparison then differs by 3e-9. So the solution is checked
    on the library's own accepted grid, against the oracle's clean backward
    Euler over exactly those times (ref_tran_grid), to 1e-9."""
    import krets_amd as kg
    say(f"parasitics LTE tps={tps}")
    nets = rectifier_batch(8, seed=9)
    cs, ctx = load(kg, nets)
    cfg = kg.SolverConfig(compat=False, voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12,
                          threads_per_system=tps)
    args = (100e-9, 0.01e-9, 1e-16, 1e-9, 1e-3, 1e-6)
    t, xs, npts, nt, rej = ctx.transient_lte(*args, 20000, cfg)
    ocfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    for i in range(8):
        oc = oracle.Circuit(nets[i], diode_parasitics=True)
        tr, xr, ntr, rejr = oc.tran_lte(*args, cfg=ocfg)
        assert npts[i] == len(tr) and nt[i] == ntr and rej[i] == rejr, (i, npts[i], len(tr), nt[i], ntr)
        tg = np.asarray(t[i][:npts[i]])
        assert np.max(np.abs(tg - tr)) <= 1e-6 * args[0], i
        xg, _ = oc.tran_grid(tg, cfg=ocfg)
        e = max(nerr(xs[i][j], xg[j]) for j in range(npts[i]))
        assert e < 1e-9, (i, e)


def test_compat_transient_ignores_junction_charge(oracle):
    """krets' own integrator (BE_COMPAT) has no junction charge: with the
    flag the series resistor is there and the charge parameters are not."""
    import krets_amd as kg
    say("parasitics BE_COMPAT")
    nets = rectifier_batch(4, seed=3)
    cs, ctx = load(kg, nets)
    cfg = kg.SolverConfig(voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12)
    xs, st, (it, _) = ctx.transient(0.1e-9, 500, method=kg.BE_COMPAT, config=cfg)
    assert (st == 0).all()
    for i in range(4):
        xr, itr = oracle.Circuit(nets[i], diode_parasitics=True).tran(0.1e-9, 50e-9,
                                                                     cfg=oracle.config(vabstol=1e-9, iabstol=1e-12))
        assert np.array_equal(it[i], itr)
        assert max(nerr(xs[i][k], xr[k]) for k in range(xr.shape[0])) < 1e-9
