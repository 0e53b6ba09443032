"""MI355X (gfx950) MNA solve backend for krets: HIP kernels behind a C ABI.

Entry points: `Solver` / `Circuit` / analyses (the krets-shaped interface),
`Context` (batched instances, host or device buffers). See include/krets_gpu.h.
"""
from ._lib import (BE, BE_COMPAT, FP64, MIXED, TRAP, KretsGpuError, lib)
from .solver import (Ac, Circuit, Context, Dc, Decade, Linear, Octave, Op, Solver, SolverConfig, Transient)

__all__ = ["Ac", "BE", "BE_COMPAT", "Circuit", "Context", "Dc", "Decade", "FP64", "KretsGpuError", "Linear",
           "MIXED", "Octave", "Op", "Solver", "SolverConfig", "TRAP", "Transient", "lib"]
