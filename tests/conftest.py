import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def read_circuit(rel):
    with open(os.path.join(GOLDEN, "circuits", rel)) as f:
        return f.read()


@pytest.fixture(scope="session")
def oracle():
    from oracle import refcpu
    refcpu.lib()
    return refcpu


# Circuits with the linear controlled sources (E VCVS, G VCCS, F CCCS, H CCVS;
# SPICE syntax, beyond krets' element set). Hand-solved operating points:
#   vcvs: V(out) = 10, I(E1) = -10 mA     vccs: V(out) = -2
#   cccs: I(Vs) = 1 mA, V(out) = 5        ccvs: V(out) = 2, I(H1) = -2 mA
CS_NETS = {
    "vcvs": "V1 in 0 1\nR1 in 0 1k\nE1 out 0 in 0 10\nRL out 0 1k\n",
    "vccs": "V1 in 0 2\nR1 in 0 1k\nG1 out 0 in 0 1m\nRL out 0 1k\n",
    "cccs": "V1 in 0 1\nR1 in mid 1k\nVs mid 0 0\nF1 0 out Vs 5\nRL out 0 1k\n",
    "ccvs": "V1 in 0 1\nR1 in mid 1k\nVs mid 0 0\nH1 out 0 Vs 2k\nRL out 0 1k\n",
}


def active_netlist(r1=10e3, r2=22e3, c1=4.7e-9, c2=1e-9, k=1.0, kf=2.0, gm=1e-3, rh=500.0):
    """A unity-gain Sallen-Key low-pass (E buffer) driving a current sense
    whose current is mirrored (F), converted back to a voltage (H) and
    transconducted (G): every controlled source in one linear circuit. The
    source is 0.5 V DC (operating point), a 1 V pulse train in transients and
    1 V in AC."""
    return (f"* active filter with every controlled source\n"
            f"V1 in 0 0.5 PULSE(0 1 0 1u 1u 40u 100u) AC 1\n"
            f"R1 in a {r1!r}\nR2 a b {r2!r}\nC1 a out {c1!r}\nC2 b 0 {c2!r}\nE1 out 0 b 0 {k!r}\n"
            f"Rs out s 1k\nVsense s 0 0\nF1 0 f Vsense {kf!r}\nRf f 0 2k\nCf f 0 1n\n"
            f"G1 g 0 f 0 {gm!r}\nRg g 0 1k\nH1 h 0 Vsense {rh!r}\nRh h 0 1k\n")


def time_aligned(tr, xr, t):
    """The oracle's states moved to the times t of another run of the same
    step controller, to first order: x_j + (x_j - x_{j-1}) / (t_j - t_{j-1})
    * (t - t_j). The accepted times of two runs differ by rounding (the
    controller amplifies it to ~1e-13 s); at 1 kHz and up to 5 V the source
    node then moves by ~1e-8 V, which is a time offset, not a state error."""
    out = xr.copy()
    for j in range(1, len(tr)):
        out[j] = xr[j] + (xr[j] - xr[j - 1]) / (tr[j] - tr[j - 1]) * (t[j] - tr[j])
    return out
