"""ctypes binding of the CPU oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by
bench.py's cpu_baseline leg. The product package never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK, MATRIX_BUILD, DECOMPOSITION, MAX_ITERATIONS, ELEMENT_NOT_FOUND = 0, 1, 2, 3, 4
INVALID_ELEMENT_FORMAT, UNEXPECTED, PARSE, EMPTY_NETLIST = 5, 6, 7, 8
COST_COMPAT, COST_LEAN = 0, 1
BE_COMPAT, BE, TRAP = 0, 1, 2
DECADE, OCTAVE, LINEAR = 0, 1, 2


class RefConfig(ctypes.Structure):
    _fields_ = [("reltol", ctypes.c_double), ("iabstol", ctypes.c_double),
                ("vabstol", ctypes.c_double), ("maxiter", ctypes.c_int32),
                ("cost", ctypes.c_int32)]


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle status {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librefcpu.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/librefcpu.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        P, D, I, C = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32, ctypes.c_char_p
        IP = ctypes.POINTER(ctypes.c_int32)
        CFG = ctypes.POINTER(RefConfig)
        L.ref_parse.restype = P
        L.ref_parse.argtypes = [C, IP, ctypes.c_char_p, I]
        L.ref_tran_grid.argtypes = [P, ctypes.POINTER(RefConfig), ctypes.POINTER(ctypes.c_double), I,
                                    ctypes.POINTER(ctypes.c_double), IP]
        L.ref_parse_ext.restype = P
        L.ref_parse_ext.argtypes = [C, I, IP, ctypes.c_char_p, I]
        L.ref_free.argtypes = [P]
        L.ref_num_unknowns.argtypes = [P]
        L.ref_num_elements.argtypes = [P]
        L.ref_unknown_name.restype = C
        L.ref_unknown_name.argtypes = [P, I]
        L.ref_element_info.argtypes = [P, I, ctypes.c_char_p, IP, IP, IP, D]
        L.ref_element_ctrl.argtypes = [P, I, IP, IP, IP]
        L.ref_set_value.argtypes = [P, I, ctypes.c_double]
        L.ref_last_error.restype = C
        L.ref_op.argtypes = [P, CFG, D, IP]
        L.ref_dc.argtypes = [P, CFG, C, ctypes.c_double, ctypes.c_double, ctypes.c_double, D, I, IP]
        L.ref_ac_frequencies.argtypes = [I, I, ctypes.c_double, ctypes.c_double, D, I]
        L.ref_ac.argtypes = [P, CFG, D, I, D]
        L.ref_tran.argtypes = [P, CFG, ctypes.c_double, ctypes.c_double, I, D, I, IP, IP]
        L.ref_tran_lte.argtypes = [P, CFG, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D, I, IP,
                                   IP, IP]
        L.ref_tran_lte_trace.argtypes = [P, CFG, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, D, I, IP]
        L.ref_wave_value.restype = ctypes.c_double
        L.ref_wave_value.argtypes = [I, D, ctypes.c_double]
        L.ref_parse_value.argtypes = [C, D]
        L.ref_pnjlim.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, D]
        L.ref_tran_adjoint.argtypes = [P, CFG, ctypes.c_double, I, I, D, D, D]
        L.ref_dense_solve.argtypes = [I, D, D, D]
        L.ref_dense_solve_c.argtypes = [I, D, D, D]
        L.ref_assemble_dc.argtypes = [P, D, D, D]
        L.ref_assemble_ac.argtypes = [P, D, ctypes.c_double, D, D]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def config(reltol=1e-3, iabstol=1e-12, vabstol=1e-6, maxiter=300, cost=COST_LEAN):
    return RefConfig(reltol, iabstol, vabstol, maxiter, cost)


def ac_frequencies(variation, points, fstart, fstop):
    buf = np.zeros(100000)
    n = lib().ref_ac_frequencies(variation, points, fstart, fstop, _dp(buf), len(buf))
    return buf[:n].copy()


def wave_value(kind, params, t):
    """Pulse::value_at (kind 1: v1 v2 td tr tf pw per) or Sinusoidal::value_at
    (kind 2: vo va f td theta phase_deg) of the oracle at time t."""
    p = np.zeros(7)
    p[:len(params)] = params
    return lib().ref_wave_value(kind, _dp(p), float(t))


def parse_value(tok):
    """utils.rs parse_value of the oracle; raises OracleError(PARSE) on failure."""
    v = ctypes.c_double()
    r = lib().ref_parse_value(tok.encode(), ctypes.byref(v))
    if r != OK:
        raise OracleError(r, f"invalid value {tok!r}")
    return v.value


def pnjlim(vnew, vold, vt, vcrit):
    """(limited voltage, whether it limited) of the oracle's SPICE3 pnjlim."""
    out = ctypes.c_double()
    lim = lib().ref_pnjlim(vnew, vold, vt, vcrit, ctypes.byref(out))
    return out.value, bool(lim)


def dense_solve(A, b):
    A = np.ascontiguousarray(A)
    n = A.shape[0]
    if np.iscomplexobj(A) or np.iscomplexobj(b):
        Ar = np.ascontiguousarray(np.asarray(A, complex).view(float))
        br = np.ascontiguousarray(np.asarray(b, complex).view(float))
        x = np.zeros(2 * n)
        r = lib().ref_dense_solve_c(n, _dp(Ar), _dp(br), _dp(x))
        out = x.view(complex)
    else:
        Ar = np.ascontiguousarray(A, float)
        br = np.ascontiguousarray(b, float)
        x = np.zeros(n)
        r = lib().ref_dense_solve(n, _dp(Ar), _dp(br), _dp(x))
        out = x
    if r != OK:
        raise OracleError(r, "singular")
    return out


class Circuit:
    """A parsed netlist in the oracle, numbered like krets."""

    def __init__(self, netlist, diode_parasitics=False):
        st = ctypes.c_int32(0)
        err = ctypes.create_string_buffer(512)
        self._h = lib().ref_parse_ext(netlist.encode(), 1 if diode_parasitics else 0, ctypes.byref(st), err, 512)
        if not self._h:
            raise OracleError(st.value, err.value.decode())
        self.n = lib().ref_num_unknowns(self._h)
        self.names = [lib().ref_unknown_name(self._h, i).decode() for i in range(self.n)]
        self.index = {k: i for i, k in enumerate(self.names)}
        self.n_elements = lib().ref_num_elements(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ref_free(self._h)
            self._h = None

    def element(self, e):
        t = ctypes.create_string_buffer(2)
        p, m, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        v = ctypes.c_double()
        r = lib().ref_element_info(self._h, e, t, ctypes.byref(p), ctypes.byref(m), ctypes.byref(b),
                                   ctypes.byref(v))
        if r:
            raise OracleError(r, "no element")
        return t.value.decode(), p.value, m.value, b.value, v.value

    def element_ctrl(self, e):
        """(ctrl_plus, ctrl_minus, ctrl_branch) of element e; -1 where not applicable."""
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        if lib().ref_element_ctrl(self._h, e, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)):
            raise OracleError(4, "no element")
        return a.value, b.value, c.value

    def set_value(self, e, v):
        lib().ref_set_value(self._h, e, float(v))

    def _check(self, r):
        if r != OK:
            raise OracleError(r, lib().ref_last_error().decode())

    def op(self, cfg=None):
        cfg = cfg or config()
        x = np.zeros(self.n)
        it = ctypes.c_int32()
        self._check(lib().ref_op(self._h, ctypes.byref(cfg), _dp(x), ctypes.byref(it)))
        return x, it.value

    def dc(self, element, start, stop, step, cfg=None, maxpts=100000):
        cfg = cfg or config()
        x = np.zeros((maxpts, self.n))
        npts = ctypes.c_int32()
        self._check(lib().ref_dc(self._h, ctypes.byref(cfg), element.encode(), start, stop, step,
                                 _dp(x), maxpts, ctypes.byref(npts)))
        return x[:npts.value].copy()

    def ac(self, freqs, cfg=None):
        cfg = cfg or config()
        f = np.ascontiguousarray(freqs, float)
        pos = int((f > 0).sum())
        x = np.zeros((max(pos, 1), self.n, 2))
        self._check(lib().ref_ac(self._h, ctypes.byref(cfg), _dp(f), len(f), _dp(x)))
        return x[:pos].view(complex)[..., 0]

    def tran(self, h, stop, method=BE_COMPAT, cfg=None):
        cfg = cfg or config()
        steps = int(round(stop / h))
        x = np.zeros((steps + 1, self.n))
        it = np.zeros(steps + 1, np.int32)
        npts = ctypes.c_int32()
        self._check(lib().ref_tran(self._h, ctypes.byref(cfg), h, stop, method, _dp(x), steps + 1,
                                   ctypes.byref(npts), _ip(it)))
        return x, it

    def tran_lte(self, tstop, h0, hmin, hmax, lte_reltol, lte_abstol, cfg=None, maxpts=200000):
        cfg = cfg or config()
        t = np.zeros(maxpts)
        x = np.zeros((maxpts, self.n))
        npts, nt, rej = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(lib().ref_tran_lte(self._h, ctypes.byref(cfg), tstop, h0, hmin, hmax, lte_reltol,
                                       lte_abstol, _dp(t), _dp(x), maxpts, ctypes.byref(npts),
                                       ctypes.byref(nt), ctypes.byref(rej)))
        k = npts.value
        return t[:k].copy(), x[:k].copy(), nt.value, rej.value

    def tran_grid(self, times, cfg=None):
        """Clean BE on the given grid (times[0] = 0): (x [npts, n], Newton total)."""
        cfg = cfg or config()
        t = np.ascontiguousarray(times, np.float64)
        x = np.zeros((len(t), self.n))
        nt = ctypes.c_int32()
        self._check(lib().ref_tran_grid(self._h, ctypes.byref(cfg), _dp(t), len(t), _dp(x), ctypes.byref(nt)))
        return x, nt.value

    def tran_lte_trace(self, tstop, h0, hmin, hmax, lte_reltol, lte_abstol, cfg=None, maxtr=400000):
        """Every attempted step of the LTE controller: rows (t, h, ratio,
        outcome, next h); ratio -1 without an estimate, outcome 1 accepted,
        0 rejected by the estimate, -1 rejected after a Newton failure."""
        cfg = cfg or config()
        tr = np.zeros((maxtr, 5))
        ntr = ctypes.c_int32()
        self._check(lib().ref_tran_lte_trace(self._h, ctypes.byref(cfg), tstop, h0, hmin, hmax, lte_reltol,
                                             lte_abstol, _dp(tr), maxtr, ctypes.byref(ntr)))
        return tr[:ntr.value].copy()

    def adjoint(self, h, nsteps, cvec, method=TRAP, cfg=None):
        cfg = cfg or config()
        c = np.ascontiguousarray(cvec, float)
        J = ctypes.c_double()
        g = np.zeros(self.n_elements)
        self._check(lib().ref_tran_adjoint(self._h, ctypes.byref(cfg), h, nsteps, method, _dp(c),
                                           ctypes.byref(J), _dp(g)))
        return J.value, g

    def assemble_dc(self, guess=None):
        A = np.zeros((self.n, self.n))
        b = np.zeros(self.n)
        g = None if guess is None else _dp(np.ascontiguousarray(guess, float))
        self._check(lib().ref_assemble_dc(self._h, g, _dp(A), _dp(b)))
        return A, b

    def assemble_ac(self, xop, f):
        A = np.zeros((self.n, self.n, 2))
        b = np.zeros((self.n, 2))
        self._check(lib().ref_assemble_ac(self._h, _dp(np.ascontiguousarray(xop, float)), f, _dp(A),
                                          _dp(b)))
        return A.view(complex)[..., 0], b.view(complex)[..., 0]
