"""GPU adjoint (kg_adjoint_forward / kg_adjoint_backward) against the CPU
oracle's FP64 discrete adjoint, central finite differences of the GPU forward
objective, and the torch.autograd wrapper.

Gradient bounds: max |g - g_ref| / max |g_ref| < 1e-9 (FP64 and mixed
precision with FP64 refinement); central differences with relative step 1e-6
agree to 1e-6 relative.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def kg():
    import krets_amd
    return krets_amd


def rlc_batch(n_inst, sections, seed):
    from krets_amd import workloads as W
    rng = np.random.default_rng(seed)
    nets = [W.c3_netlist(rng.uniform(10, 100, sections), rng.uniform(1e-9, 10e-9, sections),
                         rng.uniform(0.1e-12, 1e-12, sections)) for _ in range(n_inst)]
    import krets_amd as k
    cs = [k.Circuit(t) for t in nets]
    return nets, cs[0].elements, np.stack([c.params for c in cs]), cs[0].n, cs[0].names


def rel(g, gr):
    return np.max(np.abs(g - gr)) / np.max(np.abs(gr))


@pytest.mark.parametrize("tps,every", [(1, 64), (16, 64), (64, 64), (3064, 64), (3256, 64), (3512, 64),
                                       (3256, 1), (3256, 7), (64, 1)])  # 3000 + t: LDS-resident adjoint
@pytest.mark.parametrize("method", [1, 2])
def test_adjoint_rlc_vs_oracle(kg, oracle, method, tps, every):
    """every = 1 keeps every state as a checkpoint (the LDS-resident backward
    pass then recomputes nothing); 7 leaves a shorter last segment."""
    say(f"adjoint RLC method={method} tps={tps} checkpoint every {every}")
    nets, el, pr, n, names = rlc_batch(4, 6, seed=3)
    ctx = kg.Context()
    ctx.load(el, pr, n)
    c = np.zeros(n)
    c[names.index("V(n6)")] = 1.0
    c[names.index("I(L2)")] = 50.0
    h, N = 1e-11, 300
    J, g, st = ctx.adjoint(h, N, c, method=method, checkpoint_every=every,
                           config=kg.SolverConfig(compat=False, threads_per_system=tps))
    assert (st == 0).all()
    for i in range(4):
        Jr, gr = oracle.Circuit(nets[i]).adjoint(h, N, c, method=method)
        assert abs(J[i] - Jr) <= 1e-11 * abs(Jr)
        assert rel(g[i], gr) < 1e-9, (i, rel(g[i], gr))


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_rc_ladder_fd(kg, oracle, precision):
    from krets_amd import workloads as W
    say(f"adjoint RC ladder precision={precision}")
    b = W.c5_rc_ladder(3, sections=20, seed=5)
    ctx = kg.Context()
    b.load(ctx)
    n = b.n_unknowns
    c = np.zeros(n)
    c[b.names.index("V(n1)")] = 1.0  # non-degenerate companion objective
    h, N = 1e-9, 400
    J, g, st = ctx.adjoint(h, N, c, precision=precision, checkpoint_every=64)
    assert (st == 0).all()
    for i in range(3):
        Jr, gr = oracle.Circuit(b.netlists[i]).adjoint(h, N, c, method=oracle.TRAP)
        assert abs(J[i] - Jr) <= 1e-9 * abs(Jr)
        assert rel(g[i], gr) < 1e-9, (i, rel(g[i], gr))
    # central differences of the GPU forward objective (FP64)
    i = 1
    for e in (1, 2, 21, 40):
        p0 = b.params.copy()
        v = p0[i, e, 0]
        d = 1e-6 * v
        Js = []
        for sgn in (1, -1):
            p = p0.copy()
            p[i, e, 0] = v + sgn * d
            ctx.load(b.elements, p, n)
            Jp, _ = ctx.adjoint_forward(h, N, c)
            Js.append(Jp[i])
        fd = (Js[0] - Js[1]) / (2 * d)
        assert abs(fd - g[i, e]) <= 1e-6 * np.max(np.abs(g[i])), (e, fd, g[i, e])


def test_adjoint_torch_autograd(kg):
    import torch
    from krets_amd import workloads as W
    from krets_amd.autograd import TransientObjective
    say("adjoint torch autograd")
    b = W.c5_rc_ladder(2, sections=10, seed=9)
    ctx = kg.Context()
    c = np.zeros(b.n_unknowns)
    c[b.names.index("V(n3)")] = 1.0
    vals = torch.tensor(b.params[:, :, 0], dtype=torch.float64, requires_grad=True)
    J = TransientObjective.apply(vals, ctx, b.elements, b.params, b.n_unknowns, 1e-9, 200, c)
    J.sum().backward()
    ctx2 = kg.Context()
    b.load(ctx2)
    J2, g2, _ = ctx2.adjoint(1e-9, 200, c)
    assert np.allclose(J.detach().numpy(), J2, rtol=1e-14)
    assert np.allclose(vals.grad.numpy(), g2, rtol=1e-12, atol=0)


def test_adjoint_lds_resident_needs_fp64(kg):
    nets, el, pr, n, names = rlc_batch(2, 4, seed=5)
    ctx = kg.Context()
    ctx.load(el, pr, n)
    c = np.zeros(n)
    c[names.index("V(n4)")] = 1.0
    with pytest.raises(kg.KretsGpuError):
        ctx.adjoint(1e-11, 50, c, precision=kg.MIXED, config=kg.SolverConfig(compat=False, threads_per_system=3256))
