"""Host-side mirror of krets' solver interface over the C ABI.

`Solver(circuit, config).solve(analysis)` mirrors krets_solver::Solver
(crates/krets-solver/src/solver.rs:17-53) and returns krets' result shapes
(AnalysisResult, crates/krets-solver/src/lib.rs:13-30): dicts keyed by
"V(node)" / "I(elem)" plus "step", "frequency" or "time". `Context` is the
batched interface: many instances per call, numpy (host) or torch (device)
buffers.
"""
import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Union

import numpy as np

from . import _lib
from ._lib import KretsGpuError, lib, ptr


# ------------------------------------------------------------- netlists
class Circuit:
    """A parsed netlist in krets' numbering (parser.rs:27-154).

    diode_parasitics=True reads the diode models' series resistance (an
    internal node and resistor after krets' own unknowns and elements) and
    junction charge (cjo, vj, m, fc, tt), which krets parses or ignores
    without using (models/diode.rs:14-37)."""

    def __init__(self, text: str, diode_parasitics: bool = False):
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        flags = _lib.NETLIST_DIODE_PARASITICS if diode_parasitics else 0
        st = lib().kg_parse_netlist_ext(text.encode(), flags, ctypes.byref(h), err, 512)
        if st != _lib.OK:
            raise KretsGpuError(st, err.value.decode())
        try:
            n, ne = ctypes.c_int32(), ctypes.c_int32()
            lib().kg_circuit_size(h, ctypes.byref(n), ctypes.byref(ne))
            self.n, self.n_elements = n.value, ne.value
            self.names = [lib().kg_circuit_unknown_name(h, i).decode() for i in range(self.n)]
            self.element_ids = [lib().kg_circuit_element_id(h, e).decode() for e in range(self.n_elements)]
            self.elements = np.zeros(self.n_elements, _lib.ELEMENT_DTYPE)
            self.params = np.zeros((self.n_elements, _lib.NPARAM))
            lib().kg_circuit_tables(h, ctypes.c_void_p(self.elements.ctypes.data),
                                    self.params.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        finally:
            lib().kg_circuit_free(h)
        self.index_map = {k: i for i, k in enumerate(self.names)}

    @classmethod
    def from_file(cls, path):
        with open(path) as f:
            return cls(f.read())

    def element_index(self, identifier: str) -> int:
        try:
            return self.element_ids.index(identifier)
        except ValueError:
            raise KretsGpuError(_lib.ELEMENT_NOT_FOUND, f"Element '{identifier}' not found in the netlist")


# ------------------------------------------------------------- config / analyses
@dataclass
class SolverConfig:
    """crates/krets-solver/src/config.rs:4-36 (plus backend options)."""
    relative_tolerance: float = 1e-3
    current_absolute_tolerance: float = 1e-12
    voltage_absolute_tolerance: float = 1e-6
    maximum_iterations: int = 300
    compat: bool = True
    threads_per_system: int = 0
    refine_iters: int = 0

    def to_c(self):
        c = _lib.KgConfig()
        c.reltol = self.relative_tolerance
        c.iabstol = self.current_absolute_tolerance
        c.vabstol = self.voltage_absolute_tolerance
        c.maxiter = self.maximum_iterations
        c.compat = 1 if self.compat else 0
        c.threads_per_system = self.threads_per_system
        c.refine_iters = self.refine_iters
        return c


@dataclass
class Decade:
    points_per_decade: int


@dataclass
class Octave:
    points_per_octave: int


@dataclass
class Linear:
    total_points: int


@dataclass
class Op:
    pass


@dataclass
class Dc:
    element: str
    start: float
    stop: float
    step_size: float


@dataclass
class Ac:
    sweep: Union[Decade, Octave, Linear]
    fstart: float
    fstop: float

    def generate_frequencies(self) -> List[float]:
        """AcAnalysis::generate_frequencies (krets-parser/src/analyses.rs:80-158)."""
        f: List[float] = []
        fstart, fstop = float(self.fstart), float(self.fstop)
        if fstart <= 0.0 or fstop <= 0.0 or fstart > fstop:
            return f
        s = self.sweep
        if isinstance(s, Linear):
            if s.total_points == 1:
                f.append(fstart)
            elif s.total_points > 1:
                step = (fstop - fstart) / float(s.total_points - 1)
                f = [fstart + float(i) * step for i in range(s.total_points)]
        else:
            ppd = s.points_per_decade if isinstance(s, Decade) else s.points_per_octave
            if ppd == 0:
                return f
            nd = math.log10(fstop / fstart) if isinstance(s, Decade) else math.log2(fstop / fstart)
            total = int(_round_half_away(nd * float(ppd))) + 1
            factor = (10.0 if isinstance(s, Decade) else 2.0) ** (1.0 / float(ppd))
            cur = fstart
            for _ in range(total):
                if cur > fstop * (1.0 + 1e-9):
                    break
                f.append(cur)
                cur *= factor
                if cur > fstop and (not f or f[-1] < fstop):
                    f.append(fstop)
                    break
            if not f or f[-1] < fstop * (1.0 - 1e-9):
                f.append(fstop)
        return sorted(set(f))


@dataclass
class Transient:
    time_step: float
    stop_time: float


def _round_half_away(x):
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


# ------------------------------------------------------------- batched context
class Context:
    """One kg_ctx: a HIP stream on one GPU plus the loaded batch."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        st = lib().kg_create(ctypes.byref(h), device)
        if st != _lib.OK:
            raise KretsGpuError(st, "kg_create failed (no usable HIP device?)")
        self._h = h
        self.device = device
        self.n_instances = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().kg_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the module globals may already be gone; the
        # process then releases the device memory itself
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != _lib.OK:
            raise KretsGpuError(st, lib().kg_last_error(self._h).decode())

    def load(self, elements, params, n_unknowns, elem_offset=None):
        """Shared topology: elements [ne], params [ni, ne, NPARAM], n_unknowns int.
        Per-instance topologies: elements [sum ne], params [sum ne, NPARAM],
        n_unknowns [ni], elem_offset [ni+1]."""
        el = np.ascontiguousarray(elements, _lib.ELEMENT_DTYPE)
        pr = np.ascontiguousarray(params, np.float64)
        b = _lib.KgBatch()
        if elem_offset is None:
            ni = pr.shape[0]
            nu = np.array([n_unknowns], np.int32)
            b.n_instances, b.shared_topology, b.n_elements = ni, 1, el.shape[0]
            self.n_unknowns = np.full(ni, int(n_unknowns), np.int64)
        else:
            off = np.ascontiguousarray(elem_offset, np.int32)
            nu = np.ascontiguousarray(n_unknowns, np.int32)
            ni = len(nu)
            b.n_instances, b.shared_topology, b.n_elements = ni, 0, int(off[-1])
            b.elem_offset = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            self.n_unknowns = nu.astype(np.int64)
            self._keep_off = off
        b.n_unknowns = nu.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        b.elements = ctypes.c_void_p(el.ctypes.data)
        b.params = pr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._check(lib().kg_load_batch(self._h, ctypes.byref(b)))
        self._n_total = int(self.n_unknowns.sum())
        self.n_instances = ni
        self.shared = elem_offset is None
        self.n_elements = el.shape[0] if elem_offset is None else np.diff(elem_offset)

    def _split(self, flat, per):
        if self.shared:
            n = int(self.n_unknowns[0])
            return flat.reshape((self.n_instances,) + per + (n,))
        out, o = [], 0
        for n in self.n_unknowns:
            k = int(np.prod(per)) * int(n)
            out.append(flat[o:o + k].reshape(per + (int(n),)))
            o += k
        return out

    def op(self, config=None, out=None):
        cfg = (config or SolverConfig()).to_c()
        tot = int(self.n_unknowns.sum())
        x = np.zeros(tot) if out is None else out
        st = np.zeros(self.n_instances, np.int32)
        it = np.zeros(self.n_instances, np.int32)
        p, mem = ptr(x)
        r = lib().kg_op(self._h, ctypes.byref(cfg), p, ctypes.c_void_p(st.ctypes.data),
                        ctypes.c_void_p(it.ctypes.data), mem)
        return self._result(r, x, (), st, extra=it)

    def dc_sweep(self, element, start, stop, step, config=None):
        cfg = (config or SolverConfig()).to_c()
        npts = int(abs((stop - start) / step)) + 1
        x = np.zeros(int(self.n_unknowns.sum()) * npts)
        st = np.zeros(self.n_instances, np.int32)
        it = np.zeros(self.n_instances * npts, np.int32)
        n = ctypes.c_int32()
        r = lib().kg_dc_sweep(self._h, ctypes.byref(cfg), element, start, stop, step, ctypes.c_void_p(x.ctypes.data),
                              npts, ctypes.byref(n), ctypes.c_void_p(st.ctypes.data),
                              ctypes.c_void_p(it.ctypes.data), _lib.MEM_HOST)
        return self._result(r, x, (npts,), st, extra=it.reshape(self.n_instances, npts))

    def ac_sweep(self, freqs, precision=_lib.FP64, config=None, out=None):
        cfg = (config or SolverConfig()).to_c()
        f = np.ascontiguousarray(freqs, np.float64)
        nf = len(f)
        tot = self._n_total * nf * 2
        x = np.zeros(tot) if out is None else out
        st = np.zeros(self.n_instances, np.int32)
        err = np.zeros(self.n_instances)
        p, mem = ptr(x)
        r = lib().kg_ac_sweep(self._h, ctypes.byref(cfg), f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nf,
                              precision, p, ctypes.c_void_p(st.ctypes.data), ctypes.c_void_p(err.ctypes.data), mem)
        if out is not None:
            return self._raise_or(r, st), st, err
        xc = x.view(np.complex128)
        return self._result(r, xc, (nf,), st, extra=err)

    def transient(self, h, nsteps, method=_lib.TRAP, precision=_lib.FP64, config=None, out=None):
        cfg = (config or SolverConfig()).to_c()
        tot = int(self.n_unknowns.sum()) * (nsteps + 1)
        x = np.zeros(tot) if out is None else out
        st = np.zeros(self.n_instances, np.int32)
        it = np.zeros(self.n_instances * (nsteps + 1), np.int32)
        err = np.zeros(self.n_instances)
        p, mem = ptr(x)
        r = lib().kg_transient(self._h, ctypes.byref(cfg), h, nsteps, method, precision, p,
                               ctypes.c_void_p(st.ctypes.data), ctypes.c_void_p(it.ctypes.data),
                               ctypes.c_void_p(err.ctypes.data), mem)
        if out is not None:
            return self._raise_or(r, st), st, err
        return self._result(r, x, (nsteps + 1,), st, extra=(it.reshape(self.n_instances, nsteps + 1), err))

    def transient_lte(self, tstop, h0, hmin, hmax, lte_reltol, lte_abstol, max_points, config=None, device=False,
                      with_status=False):
        """LTE-controlled backward Euler. Returns (t [ni, max_points], x
        [ni, max_points, n], n_points, newton_total, rejected), and the
        per-instance statuses after them with with_status=True. device=True
        leaves every output in device memory (torch tensors on the context's
        GPU) instead of copying them to the host."""
        cfg = (config or SolverConfig(compat=False)).to_c()
        ni = self.n_instances
        if device:
            import torch
            if not self.shared:
                raise KretsGpuError(_lib.INVALID_ARGUMENT, "device=True needs a shared topology")
            dev = torch.device("cuda", self.device)
            t = torch.empty(ni * max_points, dtype=torch.float64, device=dev)
            x = torch.empty(int(self.n_unknowns.sum()) * max_points, dtype=torch.float64, device=dev)
            npts, nt, rej, st = (torch.empty(ni, dtype=torch.int32, device=dev) for _ in range(4))
            r = lib().kg_transient_lte(self._h, ctypes.byref(cfg), tstop, h0, hmin, hmax, lte_reltol, lte_abstol,
                                       max_points, ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(npts.data_ptr()), ctypes.c_void_p(nt.data_ptr()),
                                       ctypes.c_void_p(rej.data_ptr()), ctypes.c_void_p(st.data_ptr()),
                                       _lib.MEM_DEVICE)
            self._raise_or(r, st.cpu().numpy())
            out = (t.view(ni, max_points), x.view(ni, max_points, int(self.n_unknowns[0])), npts, nt, rej)
            return out + (st,) if with_status else out
        t = np.zeros(ni * max_points)
        x = np.zeros(int(self.n_unknowns.sum()) * max_points)
        npts = np.zeros(ni, np.int32)
        nt = np.zeros(ni, np.int32)
        rej = np.zeros(ni, np.int32)
        st = np.zeros(ni, np.int32)
        r = lib().kg_transient_lte(self._h, ctypes.byref(cfg), tstop, h0, hmin, hmax, lte_reltol, lte_abstol,
                                   max_points, ctypes.c_void_p(t.ctypes.data), ctypes.c_void_p(x.ctypes.data),
                                   ctypes.c_void_p(npts.ctypes.data), ctypes.c_void_p(nt.ctypes.data),
                                   ctypes.c_void_p(rej.ctypes.data), ctypes.c_void_p(st.ctypes.data), _lib.MEM_HOST)
        xs, _, _ = self._result(r, x, (max_points,), st)
        out = (t.reshape(ni, max_points), xs, npts, nt, rej)
        return out + (st,) if with_status else out

    def adjoint_forward(self, h, nsteps, cvec, method=_lib.TRAP, precision=_lib.FP64, checkpoint_every=64,
                        config=None):
        """Objective J = sum_s w_s c.x_s of a fixed-step linear transient;
        keeps the checkpoints for adjoint_backward. Returns (J, status)."""
        cfg = (config or SolverConfig(compat=False)).to_c()
        c = np.ascontiguousarray(cvec, np.float64)
        J = np.zeros(self.n_instances)
        st = np.zeros(self.n_instances, np.int32)
        r = lib().kg_adjoint_forward(self._h, ctypes.byref(cfg), h, nsteps, method, precision,
                                     c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), checkpoint_every,
                                     ctypes.c_void_p(J.ctypes.data), ctypes.c_void_p(st.ctypes.data), _lib.MEM_HOST)
        self._raise_or(r, st)
        return J, st

    def adjoint_backward(self):
        """dJ/d(value) of every element for the last adjoint_forward."""
        ne = int(self.n_elements)
        g = np.zeros(self.n_instances * ne)
        st = np.zeros(self.n_instances, np.int32)
        r = lib().kg_adjoint_backward(self._h, ctypes.c_void_p(g.ctypes.data), ctypes.c_void_p(st.ctypes.data),
                                      _lib.MEM_HOST)
        self._raise_or(r, st)
        return g.reshape(self.n_instances, ne), st

    def adjoint(self, h, nsteps, cvec, method=_lib.TRAP, precision=_lib.FP64, checkpoint_every=64, config=None):
        cfg = (config or SolverConfig(compat=False)).to_c()
        ni = self.n_instances
        c = np.ascontiguousarray(cvec, np.float64)
        J = np.zeros(ni)
        ne = int(self.n_elements)
        g = np.zeros(ni * ne)
        st = np.zeros(ni, np.int32)
        r = lib().kg_adjoint_transient(self._h, ctypes.byref(cfg), h, nsteps, method, precision,
                                       c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), checkpoint_every,
                                       ctypes.c_void_p(J.ctypes.data), ctypes.c_void_p(g.ctypes.data),
                                       ctypes.c_void_p(st.ctypes.data), _lib.MEM_HOST)
        self._raise_or(r, st)
        return J, g.reshape(ni, ne), st

    def last_autotune(self):
        """[(threads_per_system, kernel_ms)] timed by the last autotuning run."""
        cap = 64
        t = np.zeros(cap, np.int32)
        m = np.zeros(cap, np.float32)
        k = lib().kg_last_autotune(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), cap)
        return [(int(t[i]), float(m[i])) for i in range(min(max(k, 0), cap))]

    def use_graphs(self, enable=True):
        """Replay repeated AC sweeps as captured hipGraphs (default on)."""
        self._check(lib().kg_use_graphs(self._h, 1 if enable else 0))

    def profile(self, enable=True):
        """Returns (total_ms, launches, kernel) of the dominant analysis kernel
        (largest accumulated time) since the previous call, then resets and
        sets timing on/off."""
        ms, n, name = ctypes.c_double(), ctypes.c_int64(), ctypes.c_char_p()
        lib().kg_profile(self._h, 1 if enable else 0, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(name))
        return ms.value, n.value, (name.value or b"").decode()

    def profile_summary(self):
        """{kernel: (accumulated ms, launches)} of every kernel timed since
        the previous profile() call (no reset)."""
        need = lib().kg_profile_summary(self._h, None, 0)
        buf = ctypes.create_string_buffer(max(need, 0) + 1)
        lib().kg_profile_summary(self._h, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, n = line.rsplit(" ", 2)
            out[name] = (float(ms), int(n))
        return out

    def _raise_or(self, r, st):
        if r != _lib.OK and (st == _lib.OK).all():
            self._check(r)
        return r

    def _result(self, r, x, per, st, extra=None):
        self._raise_or(r, st)
        return self._split(x, per), st, extra


# ------------------------------------------------------------- krets mirror
class Solver:
    """krets_solver::Solver on the GPU backend (solver.rs:17-53)."""

    def __init__(self, circuit: Circuit, config: Optional[SolverConfig] = None, device: int = 0,
                 precision: int = _lib.FP64):
        self.circuit = circuit
        self.config = config or SolverConfig()
        self.precision = precision
        self._ctx = Context(device)
        self._ctx.load(circuit.elements, circuit.params[None], circuit.n)

    def _map(self, x, extra=None):
        d = {k: x[i] for i, k in enumerate(self.circuit.names)}
        if extra:
            d.update(extra)
        return d

    def _one(self, st, r):
        if st[0] != _lib.OK:
            raise KretsGpuError(int(st[0]), lib().kg_last_error(self._ctx._h).decode())

    def solve(self, analysis):
        ctx = self._ctx
        if isinstance(analysis, Op):
            xs, st, _ = ctx.op(self.config)
            self._one(st, None)
            return self._map(xs[0])
        if isinstance(analysis, Dc):
            e = self.circuit.element_index(analysis.element)
            xs, st, _ = ctx.dc_sweep(e, analysis.start, analysis.stop, analysis.step_size, self.config)
            self._one(st, None)
            return [self._map(v, {"step": float(i)}) for i, v in enumerate(xs[0])]
        if isinstance(analysis, Ac):
            f = [v for v in analysis.generate_frequencies() if v > 0.0]
            if not f:
                return []
            xs, st, _ = ctx.ac_sweep(f, self.precision, self.config)
            self._one(st, None)
            return [self._map(v, {"frequency": complex(fr, 0.0)}) for fr, v in zip(f, xs[0])]
        if isinstance(analysis, Transient):
            n = int(_round_half_away(analysis.stop_time / analysis.time_step))
            xs, st, _ = ctx.transient(analysis.time_step, n, _lib.BE_COMPAT, self.precision, self.config)
            self._one(st, None)
            return [self._map(v, {"time": float(i) * analysis.time_step}) for i, v in enumerate(xs[0])]
        raise TypeError(f"unknown analysis {analysis!r}")
