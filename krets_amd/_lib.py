"""ctypes binding of libkrets_gpu.so (include/krets_gpu.h).

The library is loaded from this package directory only; it must have been
built in-tree (python -m krets_amd.build). There is no CPU fallback: every
solve runs the HIP kernels.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkrets_gpu.so")

OK, MATRIX_BUILD, DECOMPOSITION, MAX_ITERATIONS, ELEMENT_NOT_FOUND = 0, 1, 2, 3, 4
INVALID_ELEMENT_FORMAT, UNEXPECTED, PARSE, EMPTY_NETLIST, INVALID_ARGUMENT, HIP_ERROR = 5, 6, 7, 8, 9, 100
R, C, L, V, I, D, E, G, F, H = range(10)  # kg_element_type
WAVE_NONE, WAVE_PULSE, WAVE_SIN = 0, 1, 2
FP64, MIXED = 0, 1
BE_COMPAT, BE, TRAP = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1
NPARAM = 10
NETLIST_DIODE_PARASITICS = 1  # kg_parse_netlist_ext flag
SPEC_OP, SPEC_DC, SPEC_AC, SPEC_TRANSIENT = 0, 1, 2, 3
SWEEP_DECADE, SWEEP_OCTAVE, SWEEP_LINEAR = 0, 1, 2

STATUS_NAMES = {0: "Ok", 1: "MatrixBuild", 2: "DecompositionFailed", 3: "MaximumIterationsExceeded",
                4: "ElementNotFound", 5: "InvalidElementFormat", 6: "Unexpected", 7: "ParseError",
                8: "EmptyNetlist", 9: "InvalidArgument", 100: "HipError"}


class KgElement(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("plus", ctypes.c_int32), ("minus", ctypes.c_int32),
                ("branch", ctypes.c_int32), ("g2", ctypes.c_int32), ("ctrl_plus", ctypes.c_int32),
                ("ctrl_minus", ctypes.c_int32), ("ctrl_branch", ctypes.c_int32)]


ELEMENT_DTYPE = np.dtype([("type", np.int32), ("plus", np.int32), ("minus", np.int32), ("branch", np.int32),
                          ("g2", np.int32), ("ctrl_plus", np.int32), ("ctrl_minus", np.int32),
                          ("ctrl_branch", np.int32)])


class KgBatch(ctypes.Structure):
    _fields_ = [("n_instances", ctypes.c_int32), ("shared_topology", ctypes.c_int32),
                ("n_elements", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("elem_offset", ctypes.POINTER(ctypes.c_int32)), ("n_unknowns", ctypes.POINTER(ctypes.c_int32)),
                ("elements", ctypes.c_void_p), ("params", ctypes.POINTER(ctypes.c_double))]


class KgConfig(ctypes.Structure):
    _fields_ = [("reltol", ctypes.c_double), ("iabstol", ctypes.c_double), ("vabstol", ctypes.c_double),
                ("maxiter", ctypes.c_int32), ("compat", ctypes.c_int32), ("threads_per_system", ctypes.c_int32),
                ("refine_iters", ctypes.c_int32)]


class KgSpec(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("circuit_path", ctypes.c_char * 1024), ("element", ctypes.c_char * 64),
                ("start", ctypes.c_double), ("stop", ctypes.c_double), ("step_size", ctypes.c_double),
                ("sweep", ctypes.c_int32), ("reserved", ctypes.c_int32), ("points", ctypes.c_int64),
                ("fstart", ctypes.c_double), ("fstop", ctypes.c_double), ("time_step", ctypes.c_double),
                ("stop_time", ctypes.c_double)]


class KretsGpuError(RuntimeError):
    """A non-OK kg_status; `code` mirrors krets_solver::Error."""

    def __init__(self, code, message):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


_LIB = None


def _torch_first():
    """PyTorch wheels bundle their own HIP and HSA runtimes next to the system
    ones that libkrets_gpu.so links. When PyTorch is loaded and its runtime
    started before this library is loaded, the library's HIP runtime binds to
    that same HSA runtime and device buffers pass freely between the two (the
    out= arguments). In the other order the runtime started second finds no
    GPU. So PyTorch, when importable, is loaded and started first."""
    try:
        import torch
    except ImportError:
        return
    if getattr(torch.version, "hip", None) and torch.cuda.is_available():
        torch.cuda.init()


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m krets_amd.build`")
        _torch_first()
        L = ctypes.CDLL(LIB_PATH)
        P, I32, D, C = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_char_p
        PI = ctypes.POINTER(ctypes.c_int32)
        PD = ctypes.POINTER(ctypes.c_double)
        CFG = ctypes.POINTER(KgConfig)
        L.kg_default_config.argtypes = [CFG]
        L.kg_create.argtypes = [ctypes.POINTER(P), I32]
        L.kg_destroy.argtypes = [P]
        L.kg_last_error.restype = C
        L.kg_last_error.argtypes = [P]
        L.kg_load_batch.argtypes = [P, ctypes.POINTER(KgBatch)]
        L.kg_op.argtypes = [P, CFG, P, P, P, I32]
        L.kg_dc_sweep.argtypes = [P, CFG, I32, D, D, D, P, I32, PI, P, P, I32]
        L.kg_ac_sweep.argtypes = [P, CFG, PD, I32, I32, P, P, P, I32]
        L.kg_ac_sweep_noalloc.argtypes = [P, CFG, PD, I32, I32, P, P, P]
        L.kg_transient.argtypes = [P, CFG, D, I32, I32, I32, P, P, P, P, I32]
        L.kg_transient_lte.argtypes = [P, CFG, D, D, D, D, D, D, I32, P, P, P, P, P, P, I32]
        L.kg_adjoint_transient.argtypes = [P, CFG, D, I32, I32, I32, PD, I32, P, P, P, I32]
        L.kg_adjoint_forward.argtypes = [P, CFG, D, I32, I32, I32, PD, I32, P, P, I32]
        L.kg_adjoint_backward.argtypes = [P, P, P, I32]
        L.kg_last_autotune.argtypes = [P, PI, ctypes.POINTER(ctypes.c_float), I32]
        L.kg_profile_summary.argtypes = [P, ctypes.c_char_p, I32]
        L.kg_profile.argtypes = [P, I32, PD, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_char_p)]
        L.kg_plan_export.argtypes = [P, I32, I32, I32, I32, P, D, P, I32, P, I32]
        L.kg_use_graphs.argtypes = [P, I32]
        L.kg_jit_ac_source.argtypes = [P, I32, I32, P, D, I32, I32, ctypes.c_char_p, I32]
        L.kg_jit_lte_source.argtypes = [P, I32, I32, P, D, ctypes.c_char_p, I32]
        L.kg_jit_dc_source.argtypes = [P, I32, I32, P, ctypes.c_char_p, I32]
        L.kg_jit_tran_source.argtypes = [P, I32, I32, P, D, I32, ctypes.c_char_p, I32]
        L.kg_parse_netlist.argtypes = [C, ctypes.POINTER(P), ctypes.c_char_p, I32]
        L.kg_parse_netlist_ext.argtypes = [C, I32, ctypes.POINTER(P), ctypes.c_char_p, I32]
        L.kg_circuit_free.argtypes = [P]
        L.kg_circuit_size.argtypes = [P, PI, PI]
        L.kg_circuit_unknown_name.restype = C
        L.kg_circuit_unknown_name.argtypes = [P, I32]
        L.kg_circuit_element_id.restype = C
        L.kg_circuit_element_id.argtypes = [P, I32]
        L.kg_circuit_tables.argtypes = [P, P, PD]
        L.kg_parse_spec.argtypes = [C, ctypes.POINTER(KgSpec), ctypes.c_char_p, I32]
        L.kg_parse_analysis.argtypes = [C, ctypes.POINTER(KgSpec), ctypes.c_char_p, I32]
        _LIB = L
    return _LIB


def default_config():
    c = KgConfig()
    lib().kg_default_config(ctypes.byref(c))
    return c


def ptr(a):
    """Host numpy array or torch tensor (host/device) -> (void*, mem kind)."""
    if a is None:
        return None, MEM_HOST
    if isinstance(a, np.ndarray):
        return ctypes.c_void_p(a.ctypes.data), MEM_HOST
    # torch tensor
    if getattr(a, "is_cuda", False):
        return ctypes.c_void_p(a.data_ptr()), MEM_DEVICE
    return ctypes.c_void_p(a.data_ptr()), MEM_HOST
