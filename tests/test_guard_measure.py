"""The componentwise backward-error measure of the AC accuracy guard (CPU).

The level-scheduled AC kernels judge an FP64 solve by
omega = max_i |b - A x|_i / ((|A| |x| + |b|)_i + u ||b||_inf), |z| = |re| + |im|.
The floor u ||b||_inf is there for solutions that decay towards underflow: a
2500-section RC low-pass at 100 kHz attenuates its node voltages into the
subnormal range, where the plain ratio |r|_i / (|A| |x| + |b|)_i of a solve
by the oracle's partial-pivoting LU is rounding noise of order 1. These
checks run the measure on the oracle's solutions (assembled dense by
ref_assemble_ac): with the floor it stays at rounding level at every
frequency, and where nothing underflows it stays within a factor 2 of the
plain ratio (never above it).
"""
import numpy as np

U = 2.0 ** -53


def c1(z):
    return np.abs(z.real) + np.abs(z.imag)


def measures(oc, xop, f):
    A, b = oc.assemble_ac(xop, f)
    x = oc.ac(np.array([f]))[0]
    r, d = c1(b - A @ x), c1(A) @ c1(x) + c1(b)
    plain = np.max(np.where(r > 0, r / np.where(d > 0, d, 1.0), 0.0))
    floored = np.max(r / (d + U * np.max(c1(b))))
    return plain, floored, np.min(np.abs(x[x != 0]))


def test_floor_keeps_underflowing_rows_measurable(oracle):
    from krets_amd import workloads as W
    rng = np.random.default_rng(21)
    text = W.c2_netlist(1e3 * np.exp(0.05 * rng.standard_normal(2500)),
                        1e-9 * np.exp(0.1 * rng.standard_normal(2500)))
    oc = oracle.Circuit(text)
    xop, _ = oc.op()
    for f in (1.0, 1e3):  # nothing underflows: the floor changes little
        plain, floored, xmin = measures(oc, xop, f)
        assert xmin > 1e-300 and plain < 64 * U and 0.5 * plain <= floored <= plain
    for f in (1e5, 1e6):  # subnormal node voltages: the plain ratio is noise, the floored one is not
        plain, floored, xmin = measures(oc, xop, f)
        assert xmin < 2.2250738585072014e-308 and plain > 0.1 and floored < 64 * U
