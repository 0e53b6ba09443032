"""GPU: the static-pivot guard of the FP64 transient factorisations.

Pivots are fixed once per topology from instance 0 at the call's scale
(alpha / h). A batch of three RLC ladders shares one topology: instance 0
nominal, instance 1 with L and C at 1e-2 of nominal (bound about 1e-11,
still backward stable), instance 2 at 1e-6 (bound about 1e-7, growth above
1e7).
Every FP64 mapping, the level-scheduled kernel (64 threads per system), the
LDS-resident linear kernel (3064) and the run-time specialised lane kernel
(2001), must report instance 2 as a failed decomposition and keep the other
two, with the error figure equal to the bound guard_bound_coeff(m) *
||L||_inf ||U||_inf / ||A||_inf recomputed in numpy for that mapping's pivot
order (tests/plan_emulator.py static_pivot_bound; to 1e-9 relative, the
rounding of two eliminations) and the states equal to the oracle's to 1e-9.
"""
import sys

import numpy as np
import pytest

import plan_emulator as pe
from krets_amd._lib import DECOMPOSITION
from test_plan import values

pytestmark = pytest.mark.gpu

R = [20.0, 35.0, 50.0, 25.0]
LV = [2e-9, 3e-9, 4e-9, 2.5e-9]
CV = [0.3e-12, 0.5e-12, 0.7e-12, 0.4e-12]
SCALES = [1.0, 1e-2, 1e-6]
H, N = 1e-11, 50


def say(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def nerr(x, ref):
    d = np.linalg.norm(ref)
    return np.linalg.norm(np.asarray(x) - ref) / (d if d > 0 else 1.0)


def netlists():
    from krets_amd import workloads as W
    return [W.c3_netlist(R, [v * s for v in LV], [v * s for v in CV]) for s in SCALES]


def load(kg):
    nets = netlists()
    cs = [kg.Circuit(t) for t in nets]
    ctx = kg.Context()
    ctx.load(cs[0].elements, np.stack([c.params for c in cs]), cs[0].n)
    return nets, cs, ctx


def bounds(cs, ordering, alpha):
    plan = pe.Plan(cs[0].elements, cs[0].n, pe.TRAN, ordering, cs[0].params, s_rep=alpha / H)
    return [pe.static_pivot_bound(plan, values(c, pe.TRAN, h=H, alpha=alpha)[0]) for c in cs]


@pytest.mark.parametrize("tps,ordering", [(64, pe.NESTED), (3064, pe.NESTED), (2001, pe.MINDEG)])
@pytest.mark.parametrize("method", ["BE", "TRAP"])
def test_static_pivot_guard_fixed_step(oracle, tps, ordering, method):
    import krets_amd as kg
    say(f"static-pivot guard, fixed step, {method}, tps={tps}")
    nets, cs, ctx = load(kg)
    alpha = 2.0 if method == "TRAP" else 1.0
    ref = bounds(cs, ordering, alpha)
    assert ref[0] < 1e-13 and ref[1] < 1e-10 and ref[2] > 2.0 ** -30, ref
    xs, st, (it, err) = ctx.transient(H, N, method=getattr(kg, method),
                                      config=kg.SolverConfig(compat=False, threads_per_system=tps))
    assert list(st) == [0, 0, DECOMPOSITION], st
    for k in (0, 1):
        assert abs(err[k] / ref[k] - 1.0) < 1e-9, (k, err[k], ref[k])
        xr, _ = oracle.Circuit(nets[k]).tran(H, N * H, method=getattr(oracle, method))
        assert max(nerr(xs[k][j], xr[j]) for j in range(1, N + 1)) < 1e-9, k


@pytest.mark.parametrize("tps", [64, 2001])
def test_static_pivot_guard_lte(oracle, tps):
    """The LTE-controlled transient (step sizes of its own, pivots at alpha /
    h0): the same verdicts, statuses only (the call has no error figure)."""
    import krets_amd as kg
    say(f"static-pivot guard, LTE, tps={tps}")
    nets, cs, ctx = load(kg)
    args = (2e-9, H, 1e-16, 1e-10, 1e-3, 1e-6)
    t, xs, npts, nt, rej, st = ctx.transient_lte(*args, 4000, kg.SolverConfig(compat=False, threads_per_system=tps),
                                                  with_status=True)
    assert list(st) == [0, 0, DECOMPOSITION], st
    for k in (0, 1):
        tr, xr, ntr, rejr = oracle.Circuit(nets[k]).tran_lte(*args)
        assert npts[k] == len(tr) and nt[k] == ntr and rej[k] == rejr, k
        assert max(nerr(xs[k][j], xr[j]) for j in range(npts[k])) < 1e-9, k
