"""krets.toml analysis specs and the end-to-end run of one spec.

`AnalysisSpec` mirrors krets_parser::analyses::AnalysisSpec
(crates/krets-parser/src/analyses.rs:8-23); the TOML text is read by the
library's C++ reader (kg_parse_spec). `run()` follows krets-cli's main
(crates/krets-cli/src/main.rs:33-126) without the GUI: resolve the circuit
path against the spec's folder, parse the netlist, solve on the GPU and write
`result.parquet` next to the spec.
"""
import ctypes
import os
from dataclasses import dataclass
from typing import Union

from . import _lib
from ._lib import KgSpec, KretsGpuError, lib
from .solver import Ac, Circuit, Dc, Decade, Linear, Octave, Op, Solver, SolverConfig, Transient

AnalysisT = Union[Op, Dc, Ac, Transient]


def _analysis(s: KgSpec) -> AnalysisT:
    if s.kind == _lib.SPEC_OP:
        return Op()
    if s.kind == _lib.SPEC_DC:
        return Dc(s.element.decode(), s.start, s.stop, s.step_size)
    if s.kind == _lib.SPEC_AC:
        sweep = {_lib.SWEEP_DECADE: Decade, _lib.SWEEP_OCTAVE: Octave, _lib.SWEEP_LINEAR: Linear}[s.sweep]
        return Ac(sweep(int(s.points)), s.fstart, s.fstop)
    return Transient(s.time_step, s.stop_time)


def _parse(fn, text: str) -> KgSpec:
    s = KgSpec()
    err = ctypes.create_string_buffer(512)
    st = fn(text.encode(), ctypes.byref(s), err, 512)
    if st != _lib.OK:
        raise KretsGpuError(st, err.value.decode())
    return s


def parse_analysis(text: str) -> AnalysisT:
    """A bare Analysis document (`[transient]`, `[ac]`, `[dc]` table)."""
    return _analysis(_parse(lib().kg_parse_analysis, text))


@dataclass
class AnalysisSpec:
    circuit_path: str
    analysis: AnalysisT

    @classmethod
    def from_str(cls, text: str) -> "AnalysisSpec":
        s = _parse(lib().kg_parse_spec, text)
        return cls(s.circuit_path.decode(), _analysis(s))

    @classmethod
    def from_file(cls, path) -> "AnalysisSpec":
        """AnalysisSpec::from_file (analyses.rs:19-23)."""
        with open(path) as f:
            return cls.from_str(f.read())


def resolve_circuit_path(krets_file, circuit_path) -> str:
    """main.rs:38-62: relative to the spec's folder first, else an existing
    absolute path."""
    parent = os.path.dirname(os.fspath(krets_file)) or "."
    rel = os.path.join(parent, circuit_path)
    if os.path.exists(rel):
        return rel
    if os.path.isabs(circuit_path) and os.path.exists(circuit_path):
        return circuit_path
    raise FileNotFoundError(f"Circuit file not found. Tried: relative to krets file: {rel}; "
                            f"as given: {circuit_path}")


def run(krets_file, config=None, precision=_lib.FP64, output=None):
    """Runs one krets.toml spec end to end; returns (result, parquet path).
    The result has krets' AnalysisResult shape (a map, or a list of maps)."""
    from .results import write_results_to_parquet
    spec = AnalysisSpec.from_file(krets_file)
    circuit = Circuit.from_file(resolve_circuit_path(krets_file, spec.circuit_path))
    result = Solver(circuit, config or SolverConfig(), precision=precision).solve(spec.analysis)
    out = output or os.path.join(os.path.dirname(os.fspath(krets_file)) or ".", "result.parquet")
    path = write_results_to_parquet(spec.analysis, result, out)
    return result, path
