"""Measurement modes of bench.py for the other workloads (--config c1|c3|c4|c5).

Each function returns one JSON-able dict: throughput of the batched GPU paths,
the plain FP64 GPU path where it applies, errors against the CPU oracle on a
sample, the dominant kernel's average duration and its HBM roofline share.
"""
import sys
import time

import numpy as np
import torch

import krets_amd as kg
from krets_amd import workloads as W

HBM_PEAK_GBS = 8000.0


_T0 = time.perf_counter()


def _say(msg):
    """Progress line on stderr (long runs must keep printing)."""
    print(f"[{time.perf_counter() - _T0:8.2f}s] {msg}", file=sys.stderr, flush=True)


def _oracle():
    try:
        from oracle import refcpu
        refcpu.lib()
        return refcpu
    except Exception:
        return None


def _nerr(x, r):
    d = np.linalg.norm(r)
    return float(np.linalg.norm(np.asarray(x) - r) / (d if d > 0 else 1.0))


def _werr(xs, xr):
    """Worst normwise error over the points of one trajectory, with where it occurs."""
    errs = [_nerr(xs[j], xr[j]) for j in range(len(xr))]
    j = int(np.argmax(errs))
    return errs[j], {"point": j, "err": errs[j], "norm_ref": float(np.linalg.norm(xr[j])),
                     "max_abs_diff": float(np.max(np.abs(np.asarray(xs[j]) - xr[j])))}


def _timed(ctx, fn, reps, kernels=None):
    """Wall seconds per call, the dominant kernel's mean ms per launch and
    name, and fn's result; kernels (a dict) receives every kernel's
    accumulated ms per call."""
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        res = fn()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    if kernels is not None:
        kernels.update({k: v[0] / reps for k, v in ctx.profile_summary().items()})
    kms, nl, name = ctx.profile(False)
    return dt, kms / max(nl, 1), name, res


def c1(args):
    """Plain FP64 DC operating point of the 1000-section series+shunt ladder."""
    text = W.c1_ladder()
    c = kg.Circuit(text)
    ctx = kg.Context()
    ctx.load(c.elements, c.params[None], c.n)
    ctx.op()
    dt, kms, kname, (xs, st, it) = _timed(ctx, lambda: ctx.op(), 20)
    x = xs[0]
    out = {"config": "c1 series+shunt ladder OP", "n": c.n, "solves_per_s": 1.0 / dt, "ms_per_solve": 1e3 * dt,
           "kernel": kname, "avg_kernel_ms": kms}
    phi = (1 + 5 ** 0.5) / 2
    out["I(V1)_vs_closed_form_rel"] = abs(x[c.index_map["I(V1)"]] / (-phi * 1e-3) - 1)
    R = _oracle()
    if R:
        oc = R.Circuit(text)
        xr, _ = oc.op()
        out["normwise_err_vs_oracle"] = _nerr(x, xr)
        cfg = R.config(cost=R.COST_COMPAT)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 5.0:
            oc.op(cfg)
            k += 1
        out["cpu_baseline"] = {"value": k / (time.perf_counter() - t0), "unit": "solves/s", "cores": 1,
                               "kind": "port", "sample": f"{k} OP solves, krets-cost restatement"}
        # the lean restatement: the same sparse LU from integer indices with the ordering reused
        cfg = R.config(cost=R.COST_LEAN)
        oc.op(cfg)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 3.0:
            oc.op(cfg)
            k += 1
        lean = k / (time.perf_counter() - t0)
        out["cpu_baseline"]["lean_1_thread"] = {"value": lean, "sample": f"{k} OP solves, lean restatement"}
        out["gpu_over_lean_cpu"] = (1.0 / dt) / lean
    bps = 8.0 * c.n + 8.0 * len(c.elements)  # solution written + element values read
    out["roofline"] = {"bound": "hbm", "bytes_per_solve": bps, "achieved": bps / (kms / 1e3) / 1e9,
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": bps / (kms / 1e3) / 1e9 / HBM_PEAK_GBS,
                       "note": "single 1002-unknown system: latency-bound level-scheduled LU"}
    return out


def _slice(flat, offs, i, nrow, n):
    return flat[offs[i]:offs[i] + nrow * n].view(nrow, n)


def c3(args):
    """Batched trapezoidal transient of 2048 uneven RLC ladders, 2000 steps."""
    ni = getattr(args, "c3_instances", 2048)
    N = getattr(args, "c3_steps", 2000)
    b = W.c3_rlc_uneven(ni, seed=3)
    _say(f"c3: {ni} instances generated")
    ctx = kg.Context()
    b.load(ctx)
    _say("c3: batch loaded")
    h = 10e-12
    nu = np.asarray(b.n_unknowns, np.int64)
    offs = np.concatenate([[0], np.cumsum(nu * (N + 1))])
    tot = int(offs[-1])
    dev = torch.device("cuda:0")
    buf = torch.empty(tot, dtype=torch.float64, device=dev)
    res = {"config": f"c3 {ni} uneven RLC ladders, TRAP h=10ps x {N}", "sum_unknowns": int(nu.sum())}
    sample = [0, 7, ni - 1]
    ref = {}
    for prec, key in ((kg.FP64, "fp64"), (kg.MIXED, "mixed")):
        _say(f"c3 {key}: first call (plans + autotune)")
        ctx.transient(h, N, method=kg.TRAP, precision=prec, out=buf)  # autotune + warm
        trials = ctx.last_autotune()
        _say(f"c3 {key}: autotune {trials}")
        dt, kms, kname, (r, st, err) = _timed(ctx, lambda: ctx.transient(h, N, method=kg.TRAP, precision=prec,
                                                                         out=buf), 1)
        assert (st == 0).all()
        solves = int(len(nu)) * N
        bytes_ = 8.0 * float(nu.sum()) * (N + 1)
        res[key] = {"solves_per_s": solves / dt, "ms": 1e3 * dt, "kernel": kname, "avg_kernel_ms": kms,
                    "autotune": trials,
                    "roofline": {"bound": "hbm", "achieved": bytes_ / (kms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": bytes_ / (kms / 1e3) / 1e9 / HBM_PEAK_GBS,
                                 "traffic_note": "compulsory: every unknown of every step written once"}}
        got = {i: _slice(buf, offs, i, N + 1, int(nu[i])).cpu().numpy() for i in sample}
        if prec == kg.FP64:
            ref = got
        else:
            res["mixed"]["max_normwise_err_vs_fp64_gpu"] = max(
                max(_nerr(got[i][s], ref[i][s]) for s in range(1, N + 1)) for i in sample)
    R = _oracle()
    if R:
        e = 0.0
        for i in sample[:2]:
            xr, _ = R.Circuit(b.netlists[i]).tran(h, N * h, method=R.TRAP)
            e = max(e, max(_nerr(ref[i][s], xr[s]) for s in range(1, N + 1)))
        res["fp64_normwise_err_vs_oracle"] = e
        _say("c3: oracle comparison done")
    # plain FP64 path: one instance per call
    pc = kg.Context()
    t = 0.0
    for i in sample:
        o = b.elem_offset
        pc.load(b.elements[o[i]:o[i + 1]], b.params[o[i]:o[i + 1]], [b.n_unknowns[i]], [0, o[i + 1] - o[i]])
        pc.transient(h, N, method=kg.TRAP, config=kg.SolverConfig(compat=False, threads_per_system=64))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pc.transient(h, N, method=kg.TRAP, config=kg.SolverConfig(compat=False, threads_per_system=64))
        t += time.perf_counter() - t0
        _say(f"c3: plain path instance {i}")
    res["plain_fp64_solves_per_s"] = len(sample) * N / t
    return res


def c4(args):
    """1024 half-wave rectifiers: LTE-controlled backward Euler over 5 ms."""
    b = W.c4_rectifier(1024, seed=4)
    ctx = kg.Context()
    b.load(ctx)
    cfg = kg.SolverConfig(compat=False, voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12)
    a = (5e-3, 1e-6, 1e-14, 5e-5, 1e-3, 1e-6)
    ctx.transient_lte(*a, 4000, cfg)
    dt, kms, kname, (t, xs, npts, nt, rej) = _timed(ctx, lambda: ctx.transient_lte(*a, 4000, cfg), 3)
    # the same call with every output left in HBM (no host copy of the 130 MB of states and times)
    dtd, kmsd, _, _ = _timed(ctx, lambda: ctx.transient_lte(*a, 4000, cfg, device=True), 3)
    res = {"config": "c4 1024 rectifiers, BE + LTE over 5 ms", "s": dt, "s_outputs_in_hbm": dtd,
           "solves_per_s_outputs_in_hbm": float(nt.sum()) / dtd, "kernel": kname, "avg_kernel_ms": kms,
           "accepted_steps_total": int(npts.sum()), "newton_iterations_total": int(nt.sum()),
           "rejected_total": int(rej.sum()), "newton_per_step_mean": float(nt.sum() / max(npts.sum(), 1)),
           "newton_per_step_min_instance": float((nt / np.maximum(npts, 1)).min()),
           "newton_per_step_max_instance": float((nt / np.maximum(npts, 1)).max()),
           "solves_per_s": float(nt.sum()) / dt}
    R = _oracle()
    if R:
        ok = 0
        e = 0.0
        where = []
        for i in (0, 511, 1023):
            tr, xr, ntr, rejr = R.Circuit(b.netlists[i]).tran_lte(*a, cfg=R.config(vabstol=1e-9, iabstol=1e-12))
            same = npts[i] == len(tr) and nt[i] == ntr and rej[i] == rejr
            ok += int(same)
            if same:
                ei, wi = _werr(xs[i], xr)
                wi["instance"] = i
                where.append(wi)
                e = max(e, ei)
        res["oracle_same_step_sequence"] = f"{ok}/3"
        res["normwise_err_vs_oracle"] = e
        # the same comparison with the oracle's states moved to this run's accepted times (first
        # order in the time offset): the grids differ by rounding, ~1e-13 s, which moves the 1 kHz
        # source node by ~1e-8 V without any error in the states themselves
        ea = 0.0
        for i in (0, 511, 1023):
            tr, xr, _, _ = R.Circuit(b.netlists[i]).tran_lte(*a, cfg=R.config(vabstol=1e-9, iabstol=1e-12))
            if npts[i] != len(tr):
                continue
            ti = t[i][:npts[i]]
            for j in range(1, len(tr)):
                xa = xr[j] + (xr[j] - xr[j - 1]) / (tr[j] - tr[j - 1]) * (ti[j] - tr[j])
                d = np.linalg.norm(xa)
                ea = max(ea, float(np.linalg.norm(xs[i][j] - xa) / (d if d > 0 else 1.0)))
        res["normwise_err_vs_oracle_time_aligned"] = ea
        res["worst_points"] = where
    # fixed-step BE-compat parity run at h = 1 us
    ccfg = kg.SolverConfig(voltage_absolute_tolerance=1e-9, current_absolute_tolerance=1e-12)
    ctx.transient(1e-6, 5000, method=kg.BE_COMPAT, config=ccfg)  # autotune + warm
    dt2, kms2, kname2, (xs2, st2, (it2, _)) = _timed(ctx, lambda: ctx.transient(1e-6, 5000, method=kg.BE_COMPAT,
                                                                                config=ccfg), 1)
    res["be_compat_h1us"] = {"s": dt2, "kernel": kname2, "avg_kernel_ms": kms2,
                             "newton_iterations_total": int(it2.sum()),
                             "solves_per_s": float(it2.sum()) / dt2, "all_converged": bool((st2 == 0).all()),
                             "autotune_ms_by_mapping": ctx.last_autotune()}
    if R:
        xr, itr = R.Circuit(b.netlists[3]).tran(1e-6, 5e-3, cfg=R.config(vabstol=1e-9, iabstol=1e-12))
        e2, w2 = _werr(xs2[3], xr)
        res["be_compat_h1us"]["err_vs_oracle"] = e2
        res["be_compat_h1us"]["worst_point"] = w2
        # the same with each point's norm floored at the voltage tolerance (1e-9 V): the operating point of
        # the 0 V source is exactly 0, and both codes return rounding noise of about 1e-34 there
        res["be_compat_h1us"]["err_vs_oracle_norm_floor_1e-9"] = max(
            float(np.linalg.norm(np.asarray(xs2[3][k]) - xr[k]) / max(np.linalg.norm(xr[k]), 1e-9))
            for k in range(len(xr)))
        res["be_compat_h1us"]["same_newton_counts"] = bool(np.array_equal(it2[3], itr))
    return res


def c5(args):
    """Adjoint gradients over 256 200-section RC ladders (TRAP, h = 1 ns, 2000 steps)."""
    b = W.c5_rc_ladder(256, seed=5)
    ctx = kg.Context()
    b.load(ctx)
    h, N = 1e-9, 2000
    n = b.n_unknowns
    # checkpoints every 64 steps (the workload's own setting); the backward pass recomputes all segments in
    # parallel into a trajectory of 256 x 2001 states (a few GB of the 288 GB HBM) and then sweeps it.
    # fp64_every_state keeps every state as a checkpoint instead (nothing recomputed), timed beside it
    tr_gb = 256 * (N + 1) * n * 8 / 1e9
    res = {"config": f"c5 256 RC ladders (200 sections), adjoint, TRAP h=1ns x 2000, checkpoint every 64 steps "
                     f"(segments recomputed in parallel into a {tr_gb:.1f} GB trajectory); fp64_every_state: a "
                     f"checkpoint per step"}
    R = _oracle()
    for name, node in (("spec_J_Vn200", "V(n200)"), ("companion_J_Vn1", "V(n1)")):
        c = np.zeros(n)
        c[b.names.index(node)] = 1.0
        r = {}
        for prec, key, every in ((kg.FP64, "fp64", 64), (kg.FP64, "fp64_every_state", 1), (kg.MIXED, "mixed", 64)):
            cfg = kg.SolverConfig(compat=False)
            buf = torch.empty(256 * n * (N + 1), dtype=torch.float64, device="cuda:0")
            ctx.transient(h, N, method=kg.TRAP, precision=prec, config=cfg, out=buf)  # autotune + warm
            kt, ka_all, kb_all = {}, {}, {}
            dtf, kf, _, _ = _timed(ctx, lambda: ctx.transient(h, N, method=kg.TRAP, precision=prec, config=cfg,
                                                                out=buf), 1, kt)
            del buf
            ctx.adjoint(h, N, c, precision=prec, checkpoint_every=every)
            dta, ka, _, (J, st) = _timed(ctx, lambda: ctx.adjoint_forward(h, N, c, precision=prec,
                                                                          checkpoint_every=every), 1, ka_all)
            dtb, kb, _, (g, st2) = _timed(ctx, lambda: ctx.adjoint_backward(), 1, kb_all)
            # kernel times summed over every kernel of a call (operating point, the parallel segment
            # recomputation, the sweep); the whole gradient costs the checkpointing forward plus the backward
            skt, ska, skb = sum(kt.values()), sum(ka_all.values()), sum(kb_all.values())
            r[key] = {"forward_transient_kernel_ms": skt, "adjoint_forward_kernel_ms": ska,
                      "adjoint_backward_kernel_ms": skb,
                      "adjoint_backward_kernels_ms": {k: round(v, 4) for k, v in kb_all.items()},
                      "adjoint_forward_wall_s": dta, "adjoint_backward_wall_s": dtb, "forward_transient_wall_s": dtf,
                      "backward_over_forward_wall": dtb / dtf,
                      "adjoint_total_over_forward_wall": (dta + dtb) / dtf,
                      "adjoint_total_over_forward_kernel": (ska + skb) / skt,
                      "checkpoint_every": every,
                      "J_mean": float(np.mean(J)), "J_zero_count": int((J == 0).sum())}
            if R and key == "fp64":
                # relative error over the gradient entries in the normal FP64 range: where J
                # underflows, the gradient lives in subnormals (a few significant bits each)
                errs, sub = [], 0
                for i in (0, 255):
                    Jr, gr = R.Circuit(b.netlists[i]).adjoint(h, N, c, method=R.TRAP)
                    norm = np.abs(gr) >= np.finfo(np.float64).tiny
                    sub += int(((gr != 0) & ~norm).sum())
                    if norm.any():
                        errs.append(float(np.max(np.abs(g[i] - gr)[norm]) / np.max(np.abs(gr[norm]))))
                r[key]["grad_err_vs_oracle_fp64"] = max(errs) if errs else None
                r[key]["grad_entries_subnormal_in_oracle"] = sub
            if key == "fp64":
                r["_g"] = g
            else:
                r[key]["grad_rel_diff_vs_fp64_gpu"] = float(np.max(np.abs(g - r["_g"])) / max(np.max(np.abs(r["_g"])), 1e-300))
        r.pop("_g")
        # central differences (relative step 1e-6) of the GPU FP64 forward on 4 parameters
        if name == "companion_J_Vn1":
            i = 3
            fds = []
            for e in (1, 2, 200, 399):
                p0 = b.params.copy()
                v = p0[i, e, 0]
                Js = []
                for sgn in (1, -1):
                    p = p0.copy()
                    p[i, e, 0] = v + sgn * 1e-6 * v
                    ctx.load(b.elements, p, n)
                    Jp, _ = ctx.adjoint_forward(h, N, c)
                    Js.append(Jp[i])
                fds.append((Js[0] - Js[1]) / (2e-6 * v))
            b.load(ctx)
            ctx.adjoint_forward(h, N, c)
            g, _ = ctx.adjoint_backward()
            r["fd_check"] = [{"param": e, "fd": fd, "adjoint": float(g[i, e])} for e, fd in zip((1, 2, 200, 399), fds)]
        res[name] = r
    return res


def criterion(args):
    """krets' own criterion workloads (crates/krets-solver/benches/benchmark.rs:7-111), one analysis per call
    through the krets-mirror Solver (its context persists between calls; krets builds a Solver per iteration, a
    CPU-only cost), beside the oracle in krets-cost and lean modes on one host thread."""
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "circuits")

    def circ(rel):
        with open(os.path.join(gold, rel)) as f:
            return f.read()

    work = [
        ("resistor_ladder_500", circ("resistor_ladder_500/resistor_ladder_500.cir"), kg.Op()),
        ("resistor_ladder_1000", circ("resistor_ladder_1000/resistor_ladder_1000.cir"), kg.Op()),
        ("resistor_ladder_5000", W.series_ladder(5000), kg.Op()),
        ("dc_voltage_divider_100_steps", circ("voltage_divider/voltage_divider.cir"), kg.Dc("V1", 0.0, 1.0, 0.01)),
        ("ac_low_pass_filter_100_points", circ("low_pass_filter/low_pass_filter.cir"),
         kg.Ac(kg.Linear(100), 1.0, 1000.0)),
        ("tran_dual_rc_ladder_1000_steps", circ("dual_rc_ladder/dual_rc_ladder.cir"), kg.Transient(50e-6, 50e-3)),
    ]
    R = _oracle()
    out = {}
    for name, text, an in work:
        _say(f"criterion {name}")
        r = {}
        try:
            s = kg.Solver(kg.Circuit(text))
            s.solve(an)
            torch.cuda.synchronize()
            t0, k = time.perf_counter(), 0
            while time.perf_counter() - t0 < 1.0:
                s.solve(an)
                k += 1
            r["gpu_ms_per_analysis"] = 1e3 * (time.perf_counter() - t0) / k
        except kg.KretsGpuError as e:
            r["gpu_error"] = str(e)
        if R:
            oc = R.Circuit(text)
            for mode, cost in (("cpu_krets_cost_ms", R.COST_COMPAT), ("cpu_lean_ms", R.COST_LEAN)):
                cfg = R.config(cost=cost)

                def run():
                    if isinstance(an, kg.Op):
                        oc.op(cfg)
                    elif isinstance(an, kg.Dc):
                        oc.dc(an.element, an.start, an.stop, an.step_size, cfg)
                    elif isinstance(an, kg.Ac):
                        oc.ac(np.array(an.generate_frequencies()), cfg)
                    else:
                        oc.tran(an.time_step, an.stop_time, cfg=cfg)

                run()
                t0, k = time.perf_counter(), 0
                while time.perf_counter() - t0 < 0.5:
                    run()
                    k += 1
                r[mode] = 1e3 * (time.perf_counter() - t0) / k
        out[name] = r
    return out
