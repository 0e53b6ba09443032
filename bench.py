"""Benchmark of the MNA solve path on one MI355X.

Default workload: batched AC sweep of 4096 Monte Carlo instances of a
50-section RC low-pass ladder at 200 log-spaced frequencies (1 Hz - 1 MHz):
819,200 complex 52x52 sparse solves per step. A step is one whole sweep
(operating points + all AC solves) with inputs resident in HBM and outputs
written to a device buffer.

Reported (one JSON line): batched mixed-precision (complex64 factorisation +
FP64 refinement) solves/s as `value`, next to the batched FP64 path, the plain
FP64 GPU path (one krets analysis per call, level-scheduled wave-per-system
kernel), errors against FP64 and the CPU oracle, the dominant kernel's
roofline and the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summaries (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what
    the counter reads for a device copy of known size); None if not profiled."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.txt"))):
        for line in open(path):
            m = re.match(r"traffic_bytes_per_launch\[(.+)\] = (\S+)", line.strip())
            if m and m.group(1) == kernel:
                return float(m.group(2))
    return None


def ac_bytes_per_solve(n, n_elements, nf):
    # compulsory traffic: the complex128 solution vector written per solve,
    # plus the instance's element values read once and amortised over nf
    return 16.0 * n + 8.0 * n_elements / nf


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads for the multi-core CPU baseline: the cores this process may
    run on, at most 16 (a one-GPU box's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_rate(refcpu, netlists, f, cost, threads, seconds):
    """Oracle AC solves/s over instances taken in order until `seconds` pass:
    each instance is parsed and swept over f, as krets does per analysis.
    Threads run independent instances (ctypes releases the GIL in the
    oracle's calls). Returns (solves/s, instances done)."""
    import threading
    cfg = refcpu.config(cost=cost)
    nxt = [0]
    lock = threading.Lock()
    t0 = time.perf_counter()

    def work():
        while time.perf_counter() - t0 < seconds:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= len(netlists):
                return
            refcpu.Circuit(netlists[i]).ac(f, cfg)

    ts = [threading.Thread(target=work) for _ in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    done = min(nxt[0], len(netlists))
    return done * len(f) / (time.perf_counter() - t0), done


def copy_bandwidth_gbs(torch, dev):
    """HBM bandwidth of a 1 GiB device-to-device copy (bytes read + written)."""
    src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(5):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    gbs = 5 * 2 * src.numel() * 8 / (a.elapsed_time(b) / 1e3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=6.0,
                    help="seconds of CPU work per baseline mode (compat and lean, 1 thread and all threads)")
    ap.add_argument("--no-others", action="store_true",
                    help="leave out the measurements of the other workloads (C1, C3, C4, C5) from the line")
    ap.add_argument("--plain-instances", type=int, default=8)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "criterion"],
                    help="c2 (default) prints the benchmark line; the others print measurement lines of the "
                         "other workloads")
    ap.add_argument("--c3-instances", type=int, default=2048)
    ap.add_argument("--c3-steps", type=int, default=2000)
    ap.add_argument("--probe", action="store_true",
                    help="profiling probe: one mixed and one FP64 sweep, then a 1 GiB device copy whose "
                         "counters calibrate the HBM byte counters; prints no bench line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import krets_amd as kg
    from krets_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.config != "c2":
        import bench_configs
        print(json.dumps(getattr(bench_configs, args.config)(args)), flush=True)
        return
    b = W.c2_rc_ladder_mc(args.instances, sections=50, seed=2)
    f = W.c2_frequencies(200)
    nf, n, ne, ni = len(f), b.names and len(b.names), b.elements.shape[0], b.n_instances
    nsolves = ni * nf
    ctx = kg.Context(0)
    b.load(ctx)
    out = {p: torch.empty(ni * nf * n * 2, dtype=torch.float64, device=dev) for p in (kg.FP64, kg.MIXED)}

    if args.probe:
        for p in (kg.MIXED, kg.FP64):
            ctx.ac_sweep(f, precision=p, out=out[p])
        torch.cuda.synchronize()
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "done", "copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8}))
        return

    # refinement: 0 is the library default (at most two passes; a wave stops
    # after the first once every lane's predicted error is below 1e-11),
    # > 0 a fixed number of passes
    refine = {"iters": 0}

    cfgs = {}

    def step(prec):
        key = (prec, refine["iters"])
        if key not in cfgs:
            cfgs[key] = kg.SolverConfig(refine_iters=refine["iters"]) if prec == kg.MIXED else None
        r, st, err = ctx.ac_sweep(f, precision=prec, out=out[prec], config=cfgs[key])
        if (st != 0).any():
            raise RuntimeError(f"solve failed: statuses {np.unique(st)}")
        return err

    def timed(prec, steps, warmup):
        # wall clock of the production path (repeated calls replay a captured
        # graph), then the kernel time from HIP events over separate steps
        for _ in range(warmup):
            step(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        errs = []
        for _ in range(steps):
            errs.append(step(prec))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ctx.profile(True)
        for _ in range(steps):
            step(prec)
        torch.cuda.synchronize()
        kms, nlaunch, kname = ctx.profile(False)
        return dt, kms, nlaunch, kname, np.max(errs)

    dt_d, kms_d, nl_d, kname_d, _ = timed(kg.FP64, args.steps, args.warmup)
    fp64_sps = nsolves * args.steps / dt_d
    xd = out[kg.FP64].view(ni * nf, n, 2)
    den = torch.linalg.vector_norm(xd.reshape(ni * nf, -1), dim=1)

    def mixed_err():
        xm = out[kg.MIXED].view(ni * nf, n, 2)
        num = torch.linalg.vector_norm((xm - xd).reshape(ni * nf, -1), dim=1)
        return float((num / den).max())

    # the default refinement, or else the fewest fixed passes that keep every
    # solve within 1e-9 of FP64
    tuning = ctx.last_autotune()
    for it in (0, 1, 2, 3):
        refine["iters"] = it
        step(kg.MIXED)
        if mixed_err() <= 1e-9:
            break
    tuning_mixed = ctx.last_autotune()
    dt_m, kms_m, nl_m, kname_m, est_m = timed(kg.MIXED, args.steps, args.warmup)
    mixed_sps = nsolves * args.steps / dt_m
    err_mixed_vs_fp64 = mixed_err()
    xm = out[kg.MIXED].view(ni * nf, n, 2)
    # componentwise (floor 1e-30 of the solve's max-norm)
    amax = xd.abs().amax(dim=(1, 2), keepdim=True)
    comp = ((xm - xd).abs() / torch.maximum(xd.abs(), 1e-30 * amax)).amax()
    err_comp = float(comp)

    # plain FP64 GPU path: one krets AC analysis (200 frequencies) per call,
    # level-scheduled wave-per-system kernel, one instance per context
    plain_t, plain_solves = 0.0, 0
    pctx = kg.Context(0)
    cfg = kg.SolverConfig(threads_per_system=64)
    for i in range(args.plain_instances):
        pctx.load(b.elements, b.params[i:i + 1], b.n_unknowns)
        pctx.ac_sweep(f, precision=kg.FP64, config=cfg)  # warm (symbolic pass, allocations)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pctx.ac_sweep(f, precision=kg.FP64, config=cfg)
        plain_t += time.perf_counter() - t0
        plain_solves += nf
    plain_sps = plain_solves / plain_t

    # against the CPU oracle on a sample of instances
    err_oracle = {"fp64": 0.0, "mixed": 0.0}
    cpu = None
    try:
        from oracle import refcpu
        for i in (0, ni // 2, ni - 1):
            xr = refcpu.Circuit(b.netlists[i]).ac(f)
            for key, p in (("fp64", kg.FP64), ("mixed", kg.MIXED)):
                xi = out[p].view(ni, nf, n, 2)[i].cpu().numpy()
                xi = xi[..., 0] + 1j * xi[..., 1]
                e = np.max(np.linalg.norm(xi - xr, axis=1) / np.linalg.norm(xr, axis=1))
                err_oracle[key] = max(err_oracle[key], float(e))
        if not args.no_cpu:
            nt = cpu_threads()
            rate, done = cpu_rate(refcpu, b.netlists, f, refcpu.COST_COMPAT, 1, args.cpu_seconds)
            lean, done_l = cpu_rate(refcpu, b.netlists, f, refcpu.COST_LEAN, 1, args.cpu_seconds)
            rate_t, done_t = cpu_rate(refcpu, b.netlists, f, refcpu.COST_COMPAT, nt, args.cpu_seconds)
            lean_t, done_lt = cpu_rate(refcpu, b.netlists, f, refcpu.COST_LEAN, nt, args.cpu_seconds)
            cpu = {"value": rate, "unit": "solves/s", "cores": 1, "kind": "port",
                   "sample": f"{done} instances x {nf} frequencies, krets-cost restatement (string-keyed maps, "
                             f"fresh ordering + partial-pivot sparse LU per solve), 1 thread",
                   "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                   "lean_1_thread": {"value": lean, "sample": f"{done_l} instances; the same results from integer "
                                     "indices with the ordering reused"},
                   "compat_threads": {"value": rate_t, "threads": nt, "sample": f"{done_t} instances, one per "
                                      "thread at a time"},
                   "lean_threads": {"value": lean_t, "threads": nt, "sample": f"{done_lt} instances"}}
    except Exception as e:  # the oracle is optional on the box; report why it is missing
        cpu = {"value": None, "unit": "solves/s", "cores": 1, "kind": "port", "sample": f"unavailable: {e}"}

    bps = ac_bytes_per_solve(n, ne, nf)
    avg_kernel_s = (kms_m / 1000.0) / max(nl_m, 1)
    achieved = bps * nsolves / avg_kernel_s / 1e9
    copy_gbs = copy_bandwidth_gbs(torch, dev)
    line = {
        "metric": "MNA solves/s: batched mixed-prec vs plain FP64 GPU at ≤1e-9 rel err; adj/fwd",
        "value": mixed_sps,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * dt_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 factorisation + fp64 refinement",
        "data": "synthetic (seeded lognormal R, C)",
        "config": {"workload": "batched AC sweep, 4096 Monte Carlo instances of a 50-section RC ladder x 200 "
                               "log-spaced frequencies (1 Hz-1 MHz) = 819,200 complex 52x52 solves",
                   "instances": ni, "frequencies": nf, "unknowns": n},
        "fp64_batched_solves_per_s": fp64_sps,
        "plain_fp64_solves_per_s": plain_sps,
        "ratio_mixed_vs_plain_fp64": mixed_sps / plain_sps,
        "ratio_mixed_vs_batched_fp64": mixed_sps / fp64_sps,
        "max_normwise_err_mixed_vs_fp64_gpu": err_mixed_vs_fp64,
        "max_componentwise_err_mixed_vs_fp64_gpu": err_comp,
        "max_refinement_estimate": float(est_m),
        "refinement_steps": refine["iters"] if refine["iters"] > 0 else
        "adaptive: 1 or 2 passes per wave (stop once 200*(|dx|/|x|)^2 <= 1e-11 on every lane)",
        "autotune_ms_by_threads_per_system": {"fp64": tuning, "mixed": tuning_mixed},
        "max_normwise_err_vs_oracle": err_oracle,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kname_m), "kernel": kname_m,
                     "avg_kernel_ms": 1000.0 * avg_kernel_s, "bytes_per_solve": bps,
                     "measured_copy_gbs": copy_gbs, "frac_vs_measured_copy": achieved / copy_gbs},
        "fp64_kernel": {"kernel": kname_d, "avg_kernel_ms": kms_d / max(nl_d, 1),
                        "frac": bps * nsolves / (kms_d / max(nl_d, 1) / 1e3) / 1e9 / HBM_PEAK_GBS},
        "analyses_per_s": mixed_sps / nf,
        "fp64_analyses_per_s": fp64_sps / nf,
        "cpu_baseline": cpu,
    }
    if not args.no_others:
        import bench_configs
        others = {}
        for name in ("c1", "c3", "c4", "c5", "criterion"):
            bench_configs._say(f"other workload {name}")
            try:
                others[name] = getattr(bench_configs, name)(args)
            except Exception as e:  # the headline line is still printed; the failure is reported in it
                others[name] = {"error": f"{type(e).__name__}: {e}"}
        line["other_configs"] = others
        c5 = others.get("c5", {}).get("companion_J_Vn1", {}).get("fp64", {})
        if "adjoint_total_over_forward_wall" in c5:  # (checkpointing forward + backward) / forward, FP64, C5
            line["adjoint_total_over_forward"] = c5["adjoint_total_over_forward_wall"]
            line["adjoint_total_over_forward_kernel"] = c5["adjoint_total_over_forward_kernel"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
