"""GPU: systems larger than one CU's LDS against the oracle.

A system whose values do not fit in 160 KB of LDS runs one per block with
them in a global workspace region of its block, in launches of at most 2048
systems (so the workspace stays bounded and regions are reused). Each case
first shows that its system is over the limit: with 16 threads per system
(the LDS mappings only) the library refuses it as too large for on-chip
memory. Bounds are normwise relative per solve: 1e-9 (the long ladders and
chains have condition numbers up to about 1e7, which puts rounding alone
near 1e-9 of the solution's scale).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kg():
    import krets_amd
    return krets_amd


def say(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def nerr(x, ref):
    d = np.linalg.norm(ref)
    return np.linalg.norm(np.asarray(x) - ref) / (d if d > 0 else 1.0)


def rc_ladder(rng, sections):
    from krets_amd import workloads as W
    return W.c2_netlist(1e3 * np.exp(0.05 * rng.standard_normal(sections)),
                        1e-9 * np.exp(0.1 * rng.standard_normal(sections)))


def step_ladder(rng, sections):
    from krets_amd import workloads as W
    return W.c5_netlist(1e3 * np.exp(0.05 * rng.standard_normal(sections)),
                        1e-9 * np.exp(0.1 * rng.standard_normal(sections)))


def shared_ctx(kg, nets):
    cs = [kg.Circuit(t) for t in nets]
    ctx = kg.Context()
    ctx.load(cs[0].elements, np.stack([c.params for c in cs]), cs[0].n)
    return ctx


def test_dc_sweep_larger_than_lds(kg, oracle):
    """krets' 5000-resistor chain swept over V1 = 0, 0.25 .. 1: V(k) scales
    with V1 as (1 - (k - 1) / 5000) V1."""
    from krets_amd import workloads as W
    say("DC sweep larger than LDS")
    text = W.series_ladder(5000)
    c = kg.Circuit(text)
    s = kg.Solver(c)
    with pytest.raises(kg.KretsGpuError, match="too large"):
        kg.Solver(c, kg.SolverConfig(threads_per_system=16)).solve(kg.Dc("V1", 0.0, 1.0, 0.25))
    rows = s.solve(kg.Dc("V1", 0.0, 1.0, 0.25))
    xr = oracle.Circuit(text).dc("V1", 0.0, 1.0, 0.25)
    assert len(rows) == 5 == xr.shape[0]
    ref = 1 - np.arange(5000) / 5000.0
    for j, r in enumerate(rows):
        v = np.array([r[f"V({k})"] for k in range(1, 5001)])
        assert np.max(np.abs(v - 0.25 * j * ref)) < 1e-9
        assert nerr(np.array([r[k] for k in c.names]), xr[j]) < 1e-9


def test_op_batch_in_chunks(kg):
    """2100 instances of a 3500-resistor chain (3501 unknowns, over the LDS
    limit), each with its own source voltage: two launches of the global
    workspace regions (2048 systems, then 52); every instance against the
    closed form V(k) = V1 (1 - (k - 1) / 3500)."""
    from krets_amd import workloads as W
    say("OP batch larger than LDS in two chunks")
    c = kg.Circuit(W.series_ladder(3500))
    ni = 2100
    p = np.repeat(c.params[None], ni, axis=0)
    v1 = 0.5 + np.arange(ni) / ni
    p[:, 0, 0] = v1  # V1's dc value
    ctx = kg.Context()
    ctx.load(c.elements, p, c.n)
    with pytest.raises(kg.KretsGpuError, match="too large"):
        ctx.op(kg.SolverConfig(threads_per_system=16))
    xs, st, _ = ctx.op()
    assert (st == 0).all()
    ks = np.array([c.index_map[f"V({k})"] for k in range(1, 3501)])
    shape = 1 - np.arange(3500) / 3500.0
    err = max(np.max(np.abs(xs[i][ks] - v1[i] * shape)) for i in range(ni))
    assert err < 1e-9, err


@pytest.mark.parametrize("precision,tps", [(0, 0), (0, 64), (0, 256), (1, 256)])
def test_ac_larger_than_lds(kg, oracle, precision, tps):
    """A 2500-section RC low-pass (2502 unknowns, about 200 KB per complex
    system) at 2100 frequencies: two launches (2048 systems, then 52). FP64
    sweeps 1 Hz to 1 MHz; mixed precision 100 kHz to 1 MHz, where the shunt
    capacitors make the matrix diagonally dominant (below, the resistor chain
    is conditioned near n^2 = 6e6 and complex64 factors no longer contract the
    refinement; the accuracy guard then reports those points, as for any
    system too large for its dense fixup)."""
    say(f"AC larger than LDS precision={precision} tps={tps}")
    rng = np.random.default_rng(21)
    text = rc_ladder(rng, 2500)
    ctx = shared_ctx(kg, [text])
    f = np.logspace(5 if precision else 0, 6, 2100)
    if tps == 0:
        with pytest.raises(kg.KretsGpuError, match="too large"):
            ctx.ac_sweep(f[:4], precision=precision, config=kg.SolverConfig(threads_per_system=16))
    xs, st, err = ctx.ac_sweep(f, precision=precision, config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all(), st
    xr = oracle.Circuit(text).ac(f)
    e = [nerr(xs[0][k], xr[k]) for k in range(len(f))]
    assert max(e) < 1e-9, (int(np.argmax(e)), max(e))


@pytest.mark.parametrize("method", ["BE", "TRAP"])
@pytest.mark.parametrize("precision,tps", [(0, 0), (0, 64), (1, 256)])
def test_transient_larger_than_lds(kg, oracle, method, precision, tps):
    """Two 2500-section RC ladders stepped from 0 to 1 V, 50 fixed steps."""
    say(f"transient larger than LDS {method} precision={precision} tps={tps}")
    rng = np.random.default_rng(23)
    nets = [step_ladder(rng, 2500) for _ in range(2)]
    ctx = shared_ctx(kg, nets)
    h, n = 1e-9, 50
    m = getattr(kg, method)
    if tps == 0:
        with pytest.raises(kg.KretsGpuError, match="too large"):
            ctx.transient(h, 2, method=m, precision=precision, config=kg.SolverConfig(threads_per_system=16))
    xs, st, _ = ctx.transient(h, n, method=m, precision=precision, config=kg.SolverConfig(threads_per_system=tps))
    assert (st == 0).all(), st
    for i in range(2):
        xr, _ = oracle.Circuit(nets[i]).tran(h, n * h, method=getattr(oracle, method))
        e = max(nerr(xs[i][k], xr[k]) for k in range(1, n + 1))
        assert e < 1e-9, (i, e)


def test_bucketed_transient_with_a_class_larger_than_lds(kg, oracle):
    """64 uneven RLC ladders of 8..1200 sections (up to 3600 unknowns): the
    autotuned batch puts the ladders over the LDS limit in a size class of
    their own with global workspace regions."""
    from krets_amd import workloads as W
    say("bucketed transient with a class larger than LDS")
    b = W.c3_rlc_uneven(64, smin=8, smax=1200, seed=29)
    ctx = kg.Context()
    b.load(ctx)
    h, n = 10e-12, 40
    xs, st, _ = ctx.transient(h, n, method=kg.TRAP)
    assert (st == 0).all(), st
    order = np.argsort(b.n_unknowns)
    for i in (order[0], order[32], order[-2], order[-1]):
        xr, _ = oracle.Circuit(b.netlists[i]).tran(h, n * h, method=oracle.TRAP)
        e = max(nerr(xs[i][k], xr[k]) for k in range(1, n + 1))
        assert e < 1e-9, (i, b.n_unknowns[i], e)


@pytest.mark.parametrize("tps", [64, 256])
def test_lte_larger_than_lds(kg, oracle, tps):
    """A 2500-section RC ladder under the step-size controller: the oracle's
    accepted points, Newton iterations and rejections, and its solution on
    them."""
    say(f"LTE transient larger than LDS tps={tps}")
    rng = np.random.default_rng(31)
    text = step_ladder(rng, 2500)
    ctx = shared_ctx(kg, [text])
    args = (200e-9, 1e-9, 1e-15, 20e-9, 1e-3, 1e-6)
    t, xs, npts, nt, rej = ctx.transient_lte(*args, 400, kg.SolverConfig(compat=False, threads_per_system=tps))
    tr, xr, ntr, rejr = oracle.Circuit(text).tran_lte(*args)
    assert npts[0] == len(tr) and nt[0] == ntr and rej[0] == rejr, (npts[0], len(tr), nt[0], ntr, rej[0], rejr)
    assert np.max(np.abs(np.asarray(t[0][:npts[0]]) - tr)) <= 1e-6 * args[0]
    e = max(nerr(xs[0][j], xr[j]) for j in range(npts[0]))
    assert e < 1e-9, e
