"""Synthetic workloads of the five benchmark configurations.

Every instance can be rendered as a krets netlist (floats in round-trip
form), so the same instance can be solved by krets-style tooling; the batch
form is the element table plus a per-instance parameter array for the C ABI.
All draws use numpy's PCG64 generator from a recorded seed.
"""
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib
from .solver import Circuit


@dataclass
class Batch:
    name: str
    elements: np.ndarray          # shared: [ne]; uneven: [sum ne]
    params: np.ndarray            # shared: [ni, ne, NPARAM]; uneven: [sum ne, NPARAM]
    n_unknowns: object            # shared: int; uneven: [ni]
    elem_offset: Optional[np.ndarray]
    netlists: List[str]           # one netlist per instance (may be generated lazily)
    names: Optional[List[str]] = None

    @property
    def n_instances(self):
        return self.params.shape[0] if self.elem_offset is None else len(self.n_unknowns)

    def load(self, ctx):
        ctx.load(self.elements, self.params, self.n_unknowns, self.elem_offset)


def _r(x):
    return repr(float(x))


def c1_ladder(sections=1000):
    """Series + shunt resistor ladder: V1 1 0 1; R{k} k k+1 1k; RS{k} k 0 1k."""
    lines = ["* series+shunt resistor ladder", "V1 1 0 1"]
    lines += [f"R{k} {k} {k + 1} 1000" for k in range(1, sections + 1)]
    lines += [f"RS{k} {k} 0 1000" for k in range(1, sections + 2)]
    return "\n".join(lines) + "\n"


def series_ladder(n):
    """The N-sized series resistor ladder of krets' circuits/resistor_ladder_N (V1 at node 1, R_k from node k
    to k + 1, the last to ground)."""
    lines = ["% N-sized resistor ladder", "V1 1 0 1"]
    lines += [f"R{k} {k} {k + 1} 1000" for k in range(1, n)] + [f"R{n} {n} 0 1000"]
    return "\n".join(lines) + "\n"


def c2_netlist(r, c):
    lines = ["* RC ladder low-pass", "V1 in 0 0 AC 1"]
    prev = "in"
    for k, (rk, ck) in enumerate(zip(r, c), start=1):
        lines.append(f"R{k} {prev} n{k} {_r(rk)}")
        lines.append(f"C{k} n{k} 0 {_r(ck)}")
        prev = f"n{k}"
    return "\n".join(lines) + "\n"


def c2_frequencies(n=200):
    """200 log-spaced points from 1 Hz to 1 MHz: f_j = 10^(6 j / 199)."""
    return 10.0 ** (6.0 * np.arange(n) / (n - 1))


def c2_rc_ladder_mc(n_instances=4096, sections=50, seed=2, sigma_r=0.05, sigma_c=0.10):
    rng = np.random.default_rng(seed)
    R = 1e3 * np.exp(sigma_r * rng.standard_normal((n_instances, sections)))
    C = 1e-9 * np.exp(sigma_c * rng.standard_normal((n_instances, sections)))
    tmpl = Circuit(c2_netlist(R[0], C[0]))
    params = np.repeat(tmpl.params[None], n_instances, axis=0)
    # element order: V1, R1, C1, R2, C2, ...
    params[:, 1::2, 0] = R
    params[:, 2::2, 0] = C
    nets = _Lazy(lambda i: c2_netlist(R[i], C[i]), n_instances)
    return Batch("rc_ladder_ac_mc", tmpl.elements, params, tmpl.n, None, nets, tmpl.names)


def c3_netlist(r, l, c, rise=1e-9, width=10e-9):
    lines = ["* RLC ladder", f"V1 n0 0 PULSE(0 1 0 {_r(rise)} {_r(rise)} {_r(width)} 1)"]
    for k, (rk, lk, ck) in enumerate(zip(r, l, c), start=1):
        lines += [f"R{k} n{k - 1} m{k} {_r(rk)}", f"L{k} m{k} n{k} {_r(lk)}", f"C{k} n{k} 0 {_r(ck)}"]
    return "\n".join(lines) + "\n"


def c3_rlc_uneven(n_instances=2048, smin=8, smax=512, seed=3):
    """RLC ladders with uniform section counts in [smin, smax]; R 10-100 ohm,
    L 1-10 nH, C 0.1-1 pF; PULSE source 0 -> 1 V, 1 ns rise/fall, 10 ns width,
    zero delay, 1 s period."""
    rng = np.random.default_rng(seed)
    S = rng.integers(smin, smax + 1, n_instances)
    vals = [(rng.uniform(10, 100, s), rng.uniform(1e-9, 10e-9, s), rng.uniform(0.1e-12, 1e-12, s)) for s in S]
    els, prs, nu, off = [], [], [], [0]
    cache = {}
    for i, s in enumerate(S):
        s = int(s)
        if s not in cache:
            cache[s] = Circuit(c3_netlist(*vals[i]))
        t = cache[s]
        p = t.params.copy()
        p[1::3, 0], p[2::3, 0], p[3::3, 0] = vals[i]
        els.append(t.elements)
        prs.append(p)
        nu.append(t.n)
        off.append(off[-1] + len(t.elements))
    nets = _Lazy(lambda i: c3_netlist(*vals[i]), n_instances)
    return Batch("rlc_ladder_uneven", np.concatenate(els), np.concatenate(prs), np.array(nu, np.int32),
                 np.array(off, np.int32), nets)


def c4_netlist(a, r, c, is_=1e-14):
    return (f"* half-wave rectifier\nV1 in 0 SIN(0 {_r(a)} 1k 0 0 0)\nD1 in out DMOD\n"
            f"R1 out 0 {_r(r)}\nC1 out 0 {_r(c)}\n.model DMOD D (is={_r(is_)} n=1)\n")


def c4_rectifier(n_instances=1024, seed=4):
    rng = np.random.default_rng(seed)
    A = rng.uniform(1, 5, n_instances)
    R = rng.uniform(1e3, 10e3, n_instances)
    Cv = rng.uniform(1e-6, 10e-6, n_instances)
    t = Circuit(c4_netlist(A[0], R[0], Cv[0]))
    p = np.repeat(t.params[None], n_instances, axis=0)
    p[:, 0, 4] = A  # SIN amplitude (params [3..] = vo va f td theta phase)
    p[:, 2, 0] = R
    p[:, 3, 0] = Cv
    nets = _Lazy(lambda i: c4_netlist(A[i], R[i], Cv[i]), n_instances)
    return Batch("half_wave_rectifier", t.elements, p, t.n, None, nets, t.names)


def c5_netlist(r, c):
    lines = ["* RC ladder, 1 V step", "V1 in 0 PULSE(0 1 0 0 0 1 1)"]
    prev = "in"
    for k, (rk, ck) in enumerate(zip(r, c), start=1):
        lines += [f"R{k} {prev} n{k} {_r(rk)}", f"C{k} n{k} 0 {_r(ck)}"]
        prev = f"n{k}"
    return "\n".join(lines) + "\n"


def c5_rc_ladder(n_instances=256, sections=200, seed=5, sigma=0.10):
    rng = np.random.default_rng(seed)
    R = 1e3 * np.exp(sigma * rng.standard_normal((n_instances, sections)))
    C = 1e-9 * np.exp(sigma * rng.standard_normal((n_instances, sections)))
    t = Circuit(c5_netlist(R[0], C[0]))
    p = np.repeat(t.params[None], n_instances, axis=0)
    p[:, 1::2, 0] = R
    p[:, 2::2, 0] = C
    nets = _Lazy(lambda i: c5_netlist(R[i], C[i]), n_instances)
    return Batch("rc_ladder_adjoint", t.elements, p, t.n, None, nets, t.names)


class _Lazy:
    """List-like view that renders instance netlists on demand."""

    def __init__(self, fn, n):
        self.fn, self.n = fn, n

    def __len__(self):
        return self.n

    def __getitem__(self, i):
        return self.fn(i)
