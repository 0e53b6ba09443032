"""GPU parity for the linear controlled sources (E VCVS, G VCCS, F CCCS,
H CCVS; SPICE syntax, the components of krets' README to-do list,
README.md:56-62).

A batch of active filters that holds all four (a Sallen-Key low-pass with an
E buffer, an F mirror of a sense current, an H transresistance and a G
transconductor), with different values per instance, runs every analysis and
every work mapping through the C ABI and is compared with the oracle, which
the CPU tests pin to closed forms. Bounds are normwise relative per solve:
1e-12 for operating points, DC sweeps and FP64 AC, 1e-9 for mixed precision,
transients and gradients.
"""
import sys

import numpy as np
import pytest

from conftest import active_netlist

pytestmark = pytest.mark.gpu

NI = 6


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


@pytest.fixture(scope="module")
def batch(kg):
    rng = np.random.default_rng(17)
    nets = [active_netlist(rng.uniform(5e3, 20e3), rng.uniform(10e3, 40e3), rng.uniform(2e-9, 8e-9),
                           rng.uniform(0.5e-9, 2e-9), rng.uniform(0.8, 1.5), rng.uniform(0.5, 4.0),
                           rng.uniform(0.5e-3, 3e-3), rng.uniform(100.0, 2e3)) for _ in range(NI)]
    cs = [kg.Circuit(t) for t in nets]
    assert all(c.names == cs[0].names for c in cs)
    return nets, cs[0].elements, np.stack([c.params for c in cs]), cs[0].n, cs[0].names


def context(kg, batch):
    ctx = kg.Context()
    ctx.load(batch[1], batch[2], batch[3])
    return ctx


def test_controlled_op_and_dc(kg, oracle, batch):
    nets, _, _, _, names = batch
    say("controlled sources: OP and DC sweep")
    ctx = context(kg, batch)
    x, st, _ = ctx.op()
    assert (st == 0).all()
    xd, std, _ = ctx.dc_sweep(0, -1.0, 1.0, 0.25)  # element 0 is V1
    assert (std == 0).all()
    for i in range(NI):
        oc = oracle.Circuit(nets[i])
        assert oc.names == names
        assert nerr(x[i], oc.op()[0]) < 1e-12
        xr = oc.dc("V1", -1.0, 1.0, 0.25)
        assert xd[i].shape == xr.shape
        assert max(nerr(xd[i][p], xr[p]) for p in range(xr.shape[0])) < 1e-12
    # single-circuit Solver path, krets-style result maps
    s = kg.Solver(kg.Circuit(nets[0]))
    r = s.solve(kg.Op())
    assert nerr([r[k] for k in names], oracle.Circuit(nets[0]).op()[0]) < 1e-12


@pytest.mark.parametrize("precision,tps", [(0, 0), (0, 1), (0, 16), (0, 1016), (0, 2001), (0, 2011),
                                           (1, 0), (1, 16), (1, 2001), (1, 2002), (1, 2012)])
def test_controlled_ac(kg, oracle, batch, precision, tps):
    nets = batch[0]
    say(f"controlled sources: AC precision={precision} tps={tps}")
    ctx = context(kg, batch)
    f = np.logspace(1, 6, 41)
    xs, st, err = ctx.ac_sweep(f, precision=precision, config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all()
    bound = 1e-12 if precision == 0 else 1e-9
    for i in range(NI):
        xr = oracle.Circuit(nets[i]).ac(f)
        e = max(nerr(xs[i][j], xr[j]) for j in range(len(f)))
        assert e < bound, (i, e)
    if precision == 1:
        assert (err <= 1e-9).all()


@pytest.mark.parametrize("tps", [0, 16, 1064, 2001, 3064, 3256])
@pytest.mark.parametrize("method", [0, 1, 2])
def test_controlled_transient(kg, oracle, batch, method, tps):
    nets = batch[0]
    say(f"controlled sources: transient method={method} tps={tps}")
    ctx = context(kg, batch)
    h, N = 2e-6, 150
    xs, st, _ = ctx.transient(h, N, method=method, config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all()
    for i in range(NI):
        xr, _ = oracle.Circuit(nets[i]).tran(h, N * h, method=method)
        e = max(nerr(xs[i][k], xr[k]) for k in range(N + 1))
        assert e < 1e-9, (i, e)


def test_controlled_transient_mixed(kg, oracle, batch):
    nets = batch[0]
    say("controlled sources: transient mixed precision")
    ctx = context(kg, batch)
    h, N = 2e-6, 150
    xs, st, (_, err) = ctx.transient(h, N, method=kg.TRAP, precision=kg.MIXED)
    assert (st == 0).all()
    for i in range(NI):
        xr, _ = oracle.Circuit(nets[i]).tran(h, N * h, method=kg.TRAP)
        assert max(nerr(xs[i][k], xr[k]) for k in range(N + 1)) < 1e-9


@pytest.mark.parametrize("tps", [0, 16, 3064, 3256])
@pytest.mark.parametrize("method", [1, 2])
def test_controlled_adjoint(kg, oracle, batch, method, tps):
    nets, el, _, n, names = batch
    say(f"controlled sources: adjoint method={method} tps={tps}")
    ctx = context(kg, batch)
    c = np.zeros(n)
    c[names.index("V(g)")] = 1.0
    c[names.index("V(out)")] = 0.5
    h, N = 2e-6, 150
    J, g, st = ctx.adjoint(h, N, c, method=method, checkpoint_every=16,
                           config=kg.SolverConfig(compat=False, threads_per_system=tps))
    assert (st == 0).all()
    gains = [e for e in range(len(el)) if el[e]["type"] >= 6]
    assert len(gains) == 4
    for i in range(NI):
        Jr, gr = oracle.Circuit(nets[i]).adjoint(h, N, c, method=method)
        assert abs(J[i] - Jr) <= 1e-11 * abs(Jr)
        assert np.max(np.abs(g[i] - gr)) < 1e-9 * np.max(np.abs(gr)), i
        assert (g[i][gains] == 0).all()


@pytest.mark.parametrize("tps", [1, 16, 2001])  # 2001: lane kernel specialised at run time
def test_controlled_lte(kg, oracle, batch, tps):
    from conftest import time_aligned
    nets = batch[0]
    say(f"controlled sources: LTE transient tps={tps}")
    ctx = context(kg, batch)
    cfg = kg.SolverConfig(compat=False, voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12,
                          threads_per_system=tps)
    args = (100e-6, 1e-8, 1e-16, 5e-6, 1e-3, 1e-6)
    t, xs, npts, nt, rej = ctx.transient_lte(*args, 2000, cfg)
    ocfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    for i in range(NI):
        tr, xr, ntr, rejr = oracle.Circuit(nets[i]).tran_lte(*args, cfg=ocfg)
        assert npts[i] == len(tr) and rej[i] == rejr and nt[i] == ntr, (i, npts[i], len(tr), rej[i], rejr)
        xa = time_aligned(tr, xr, t[i][:npts[i]])
        assert max(nerr(xs[i][j], xa[j]) for j in range(len(tr))) < 1e-9
