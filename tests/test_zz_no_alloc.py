"""GPU: the allocation-free AC sweep (kg_ac_sweep_noalloc) against the
allocating kg_ac_sweep it mirrors, through the C ABI.

Both run the same kernels on the same inputs, so results, statuses and error
figures must be identical bit for bit. Each case runs two allocation-free
calls in a row, the second with fewer frequencies, so that the second call
starts from the first call's leftovers in the shared scratch array (queue
count, error figures). Cases: the batched RC ladder in FP64 and mixed
precision, and the wide-range RLC batch whose solves go through the
accuracy guard's fixup (its queue lives in the scratch array).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from krets_amd import _lib


def say(msg):
    print(f"[gpu case] {msg}", flush=True)


def noalloc_call(kg, ctx, freqs, prec, cfg, n):
    import torch
    ni, nf = ctx.n_instances, len(freqs)
    dev = torch.device("cuda", ctx.device)
    x = torch.full((ni * nf * n * 2,), float("nan"), dtype=torch.float64, device=dev)
    st = torch.full((ni,), -1, dtype=torch.int32, device=dev)
    err = torch.full((ni,), -1.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    f = np.ascontiguousarray(freqs, np.float64)
    c = cfg.to_c()
    r = kg.lib().kg_ac_sweep_noalloc(ctx._h, ctypes.byref(c), f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nf,
                                      prec, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(st.data_ptr()),
                                      ctypes.c_void_p(err.data_ptr()))
    torch.cuda.synchronize()
    xs = x.cpu().numpy().view(np.complex128).reshape(ni, nf, n)
    return r, xs, st.cpu().numpy(), err.cpu().numpy()


def check_pair(kg, ctx, n, f_big, f_small, prec, cfg):
    ref = [ctx.ac_sweep(f, precision=prec, config=cfg) for f in (f_big, f_small)]  # allocating: plans, mapping
    got = [noalloc_call(kg, ctx, f, prec, cfg, n) for f in (f_big, f_small)]  # two in a row, smaller second
    for (xr, sr, er), (r, x, s, e) in zip(ref, got):
        assert r == _lib.OK or (sr != 0).any(), r
        assert np.array_equal(s, sr), (s, sr)
        assert np.array_equal(x.view(np.float64), np.asarray(xr).view(np.float64))
        assert np.array_equal(e, er)


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_ac_sweep_rc_batch(prec):
    import krets_amd as kg
    from krets_amd import workloads as W
    say(f"no-alloc RC batch prec={prec}")
    b = W.c2_rc_ladder_mc(130, sections=50, seed=9)
    ctx = kg.Context()
    b.load(ctx)
    f = W.c2_frequencies(200)
    fs = f[np.linspace(0, len(f) - 1, 50).round().astype(int)]  # same range: the same plans
    check_pair(kg, ctx, 52, f, fs, prec, kg.SolverConfig())


@pytest.mark.parametrize("prec,tps", [(0, 2001), (1, 2002), (0, 16)])
def test_noalloc_matches_ac_sweep_through_the_fixup(prec, tps):
    import krets_amd as kg
    from test_gpu_guard import FREQS, wide_rlc_batch
    say(f"no-alloc wide RLC prec={prec} tps={tps}")
    cs = [kg.Circuit(t) for t in wide_rlc_batch()]
    ctx = kg.Context()
    ctx.load(cs[0].elements, np.stack([c.params for c in cs]), cs[0].n)
    fs = FREQS[::2]  # 13 of the 25 points, the same end points
    check_pair(kg, ctx, cs[0].n, FREQS, fs, prec, kg.SolverConfig(threads_per_system=tps))


def test_noalloc_refusals():
    import torch
    import krets_amd as kg
    from krets_amd import workloads as W
    say("no-alloc refusals")
    b = W.c2_rc_ladder_mc(8, sections=50, seed=9)
    ctx = kg.Context()
    b.load(ctx)
    f = W.c2_frequencies(20)
    # nothing cached yet: no plan, no mapping
    r, _, _, _ = noalloc_call(kg, ctx, f, 0, kg.SolverConfig(), 52)
    assert r == _lib.INVALID_ARGUMENT and "allocation-free" in kg.lib().kg_last_error(ctx._h).decode()
    # host buffers are refused
    cfg = kg.SolverConfig().to_c()
    x = np.zeros(8 * 20 * 52 * 2)
    st = np.zeros(8, np.int32)
    r = kg.lib().kg_ac_sweep_noalloc(ctx._h, ctypes.byref(cfg), f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      20, 0, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(st.ctypes.data), None)
    assert r == _lib.INVALID_ARGUMENT
    # a workload whose operating point alone exceeds the scratch array
    big = W.c2_rc_ladder_mc(12000, sections=100, seed=1)  # 12000 x 102 unknowns x 8 bytes = 9.8 MB
    ctx2 = kg.Context()
    big.load(ctx2)
    xb = torch.empty(2, dtype=torch.float64, device="cuda")
    sb = torch.empty(12000, dtype=torch.int32, device="cuda")
    r = kg.lib().kg_ac_sweep_noalloc(ctx2._h, ctypes.byref(cfg), f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      20, 0, ctypes.c_void_p(xb.data_ptr()), ctypes.c_void_p(sb.data_ptr()), None)
    msg = kg.lib().kg_last_error(ctx2._h).decode()
    assert r == _lib.INVALID_ARGUMENT and "bytes of scratch" in msg and "8388608" in msg, msg
