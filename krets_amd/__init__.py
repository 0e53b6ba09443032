"""MI355X (gfx950) MNA solve backend for krets: HIP kernels behind a C ABI.

Entry points: `Solver` / `Circuit` / analyses (the krets-shaped interface),
`Context` (batched instances, host or device buffers), `AnalysisSpec` / `run`
(krets.toml specs end to end, parquet results). See include/krets_gpu.h.
"""
from ._lib import (BE, BE_COMPAT, FP64, MIXED, TRAP, KretsGpuError, lib)
from .solver import (Ac, Circuit, Context, Dc, Decade, Linear, Octave, Op, Solver, SolverConfig, Transient)
from .spec import AnalysisSpec, parse_analysis, run

__all__ = ["Ac", "AnalysisSpec", "BE", "BE_COMPAT", "Circuit", "Context", "Dc", "Decade", "FP64", "KretsGpuError",
           "Linear", "MIXED", "Octave", "Op", "Solver", "SolverConfig", "TRAP", "Transient", "lib", "parse_analysis",
           "run"]
