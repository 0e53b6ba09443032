"""Pins the CPU oracle (oracle/refcpu.cpp) to krets' own known answers, its
golden AC vector, closed forms and a dense long-double brute force.

Known-answer values are those asserted by crates/krets-solver/tests/*.rs and
crates/krets-parser/src/elements/voltage_source.rs; tolerances here are the
reference's own or tighter where a closed form exists.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, read_circuit


def sol(c, x):
    return {k: x[i] for k, i in c.index.items()}


@pytest.mark.parametrize("cost", [0, 1])
def test_op_farid_najm(oracle, cost):
    # crates/krets-solver/tests/test_op.rs:25-46
    c = oracle.Circuit(read_circuit("circuit_simulation_farid_n_najm/circuit_simulation_farid_n_najm.cir"))
    x, _ = c.op(oracle.config(cost=cost))
    s = sol(c, x)
    exp = {"V(4)": 1.9888, "V(8)": 1.0, "V(3)": 2.00879, "V(2)": 1.80879, "V(6)": 1.98814, "V(5)": 2.0,
           "V(1)": 1.88527, "V(7)": 3.98814, "I(V3)": 0.0, "I(V2)": -199.88e-3, "I(V1)": -198.88e-3}
    for k, v in exp.items():
        assert abs(s[k] - v) < 1e-3, k


def test_op_case_1(oracle):
    # test_op.rs:48-68 (exact integers: tightened to 1e-12)
    c = oracle.Circuit("V1 2 1 32\nR1 1 0 2\nR2 2 3 4\nR3 2 0 8\nV2 3 0 20\n")
    s = sol(c, c.op()[0])
    for k, v in {"V(1)": -8.0, "V(2)": 24.0, "V(3)": 20.0, "I(V1)": -4.0, "I(V2)": 1.0}.items():
        assert abs(s[k] - v) < 1e-12, k


def test_op_divider_filters_basic(oracle):
    c = oracle.Circuit(read_circuit("voltage_divider/voltage_divider.cir"))
    s = sol(c, c.op()[0])  # test_op.rs:70-81
    assert abs(s["V(out)"] - 2 / 3) < 1e-14 and abs(s["I(V1)"] + 1 / 3000) < 1e-15
    for f in ("low_pass_filter/low_pass_filter.cir", "high_pass_filter/high_pass_filter.cir"):
        c = oracle.Circuit(read_circuit(f))  # test_op.rs:83-105
        s = sol(c, c.op()[0])
        assert abs(s["V(in)"] - 1) < 1e-12 and abs(s["V(out)"] - 1) < 1e-12
    c = oracle.Circuit(read_circuit("basic_001/basic_001.cir"))  # test_op.rs:107-118
    s = sol(c, c.op()[0])
    assert abs(s["V(1)"] - 3) < 1e-12 and abs(s["V(2)"] - 0.5) < 1e-12 and abs(s["I(V4)"] + 0.5) < 1e-12


def test_op_series_ladders(oracle):
    # test_op.rs:120-129; closed form V(k) = 1 - (k-1)/N, I(V1) = -1/(N*1000)
    for N in (500, 1000):
        c = oracle.Circuit(read_circuit(f"resistor_ladder_{N}/resistor_ladder_{N}.cir"))
        x, _ = c.op()
        s = sol(c, x)
        ks = np.arange(1, N + 1)
        v = np.array([s[f"V({k})"] for k in ks])
        assert np.max(np.abs(v - (1 - (ks - 1) / N))) < 1e-12
        assert abs(s["I(V1)"] * N * 1000.0 + 1.0) < 1e-11


def test_series_ladder_generator_matches_krets_files():
    """workloads.series_ladder(N) (used for krets' resistor_ladder_5000, whose
    circuit file is not among the fixtures) writes the element lines of krets'
    resistor_ladder_N files, which are."""
    import re
    from krets_amd import workloads as W

    def elements(text):
        return [ln.split() for ln in text.splitlines() if re.match(r"^[RV]\d", ln)]

    for N in (500, 1000):
        assert elements(W.series_ladder(N)) == elements(read_circuit(f"resistor_ladder_{N}/resistor_ladder_{N}.cir"))


def test_op_diode(oracle):
    # test_op.rs:131-140; closed form: scalar Newton on 1 = 1000*I + V, I = Is(e^{V/Vt}-1)
    c = oracle.Circuit(read_circuit("diode_iv_curve/diode_iv_curve.cir"))
    x, it = c.op()
    s = sol(c, x)
    assert abs(s["V(out)"] - 0.517) < 1e-3
    vt = 1.380649e-23 * 300.0 / 1.602176634e-19
    v = 0.5
    for _ in range(100):
        f = (1 - v) / 1000 - 1e-12 * (math.exp(v / vt) - 1)
        df = -1 / 1000 - 1e-12 / vt * math.exp(v / vt)
        v -= f / df
    assert abs(s["V(out)"] - v) < 1e-5  # krets' Newton stops at reltol 1e-3
    assert it >= 2


def test_dc_divider_sweep(oracle):
    # test_dc.rs:25-54 (pins the I(V1) sign convention)
    c = oracle.Circuit(read_circuit("voltage_divider/voltage_divider.cir"))
    x = c.dc("V1", 0.0, 1.0, 1.0)
    assert x.shape[0] == 2
    s0, s1 = sol(c, x[0]), sol(c, x[1])
    assert abs(s0["V(in)"]) < 1e-15 and abs(s0["I(V1)"]) < 1e-15
    assert abs(s1["V(out)"] - 2 / 3) < 1e-14 and abs(s1["I(V1)"] + 1 / 3000) < 1e-16
    with pytest.raises(oracle.OracleError) as e:
        c.dc("V9", 0, 1, 1)
    assert e.value.code == oracle.ELEMENT_NOT_FOUND


def test_ac_known_answers(oracle):
    c = oracle.Circuit(read_circuit("low_pass_filter/low_pass_filter.cir"))
    x = c.ac([1000.0])[0]  # test_ac.rs:25-46
    assert abs(x[c.index["V(out)"]] - (2.470452e-02 - 1.55223e-01j)) < 1e-6
    w = 2 * math.pi * 1000 * 1e3 * 1e-6  # test_ac.rs:76-105 analytic RC
    assert abs(x[c.index["V(out)"]] - 1 / (1 + 1j * w)) < 1e-14
    c = oracle.Circuit(read_circuit("high_pass_filter/high_pass_filter.cir"))
    x = c.ac([1000.0])[0]  # test_ac.rs:48-74
    assert abs(x[c.index["V(out)"]] - (7.169568e-01 - 4.50477e-01j)) < 1e-6
    assert abs(x[c.index["I(V1)"]] - (-7.16957e-03 + 4.504772e-03j)) < 1e-8


def test_ac_golden_parquet(oracle):
    # circuits/low_pass_filter/ac/result.parquet: the only full-precision golden vector
    g = json.load(open(os.path.join(GOLDEN, "low_pass_ac_result.json")))
    cols = {k: np.array([float(v) for v in vals]) for k, vals in g["columns"].items()}
    f = oracle.ac_frequencies(oracle.LINEAR, 10, 0.1, 1000.0)
    assert np.array_equal(f, cols["frequency"])  # generate_frequencies is bit-exact
    c = oracle.Circuit(read_circuit("low_pass_filter/ac/ac.cir"))
    x = c.ac(f)
    for name in ("I(V1)", "V(in)", "V(out)"):
        v = x[:, c.index[name]]
        mag = np.sqrt(v.real ** 2 + v.imag ** 2)
        ph = np.arctan2(v.imag, v.real) * 180.0 / math.pi
        assert np.max(np.abs(mag - cols[name + "_mag"]) / np.abs(cols[name + "_mag"])) < 1e-13
        assert np.max(np.abs(ph - cols[name + "_phase_deg"])) < 1e-11


def test_ac_frequencies_decade_octave(oracle):
    f = oracle.ac_frequencies(oracle.DECADE, 10, 1.0, 1e6)
    assert len(f) == 61 and f[0] == 1.0 and f[-1] == 1e6
    assert np.all(np.diff(f) > 0)
    f = oracle.ac_frequencies(oracle.OCTAVE, 1, 1.0, 8.0)
    assert np.allclose(f, [1, 2, 4, 8])
    assert len(oracle.ac_frequencies(oracle.LINEAR, 5, 0.0, 10.0)) == 0


def test_transient_lowpass_lag(oracle):
    # test_transient.rs:180-212: the source-time-lag pin (index 42 -> 0.8647)
    c = oracle.Circuit(read_circuit("low_pass_filter/transient.cir"))
    x, _ = c.tran(50e-6, 20e-3)
    vo = x[:, c.index["V(out)"]]
    assert abs(vo[0]) < 1e-12 and abs(vo[42] - 0.8647) < 1e-3 and abs(vo[-1] - 1.0) < 1e-3
    # closed form of lagged BE: v_n = v_{n-1} + (u(t_{n-1}) - v_{n-1}) * a, a = h/(RC + h)
    h, rc = 50e-6, 1e-3
    a = h / (rc + h)
    v, ref = 0.0, [0.0]
    for n in range(1, len(vo)):
        tp = (n - 1) * h
        u = 0.0 if tp < 1e-6 else (min((tp - 1e-6) / 1e-6, 1.0))
        v = (v * rc / h + u) * a
        ref.append(v)
    assert np.max(np.abs(vo - np.array(ref))) < 1e-13


def test_transient_rl_dual_rc_rectifier(oracle):
    c = oracle.Circuit(read_circuit("high_pass_filter/transient.cir"))  # test_transient.rs:214-238
    x, _ = c.tran(10e-6, 2e-3)
    assert abs(x[-1, c.index["V(out)"]] - 1.0) < 1e-3
    c = oracle.Circuit(read_circuit("dual_rc_ladder/dual_rc_ladder.cir"))  # test_transient.rs:139-161
    x, _ = c.tran(50e-6, 50e-3)
    assert abs(x[-1, c.index["V(out)"]] - 0.989) < 1e-3
    c = oracle.Circuit(read_circuit("rectifier/rectifier.cir"))  # test_transient.rs:163-178 (smoke)
    x, it = c.tran(50e-6, 50e-3)
    assert it.max() < 300 and np.all(np.isfinite(x))


def test_sparse_lu_vs_dense_long_double(oracle):
    c = oracle.Circuit(read_circuit("circuit_simulation_farid_n_najm/circuit_simulation_farid_n_najm.cir"))
    A, b = c.assemble_dc()
    x = oracle.dense_solve(A, b)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-13, atol=1e-15)
    assert np.max(np.abs(x - c.op()[0])) < 1e-14
    c = oracle.Circuit(read_circuit("high_pass_filter/high_pass_filter.cir"))
    xop, _ = c.op()
    A, b = c.assemble_ac(xop, 1234.5)
    assert np.max(np.abs(oracle.dense_solve(A, b) - c.ac([1234.5])[0])) < 1e-15


def test_errors(oracle):
    with pytest.raises(oracle.OracleError) as e:
        oracle.Circuit("")
    assert e.value.code == oracle.EMPTY_NETLIST
    c = oracle.Circuit(read_circuit("diode_iv_curve/diode_iv_curve.cir"))
    with pytest.raises(oracle.OracleError) as e:
        c.op(oracle.config(maxiter=1))
    assert e.value.code == oracle.MAX_ITERATIONS
    c = oracle.Circuit("I1 1 0 1\nR1 1 0 1\n")  # AC current source lacks (ic,ic): singular
    with pytest.raises(oracle.OracleError) as e:
        c.ac([1.0])
    assert e.value.code == oracle.DECOMPOSITION
    c = oracle.Circuit("V1 1 0 1\nR1 1 2 1\nC1 2 0 1u G2\n")  # G2 capacitor is singular in OP
    with pytest.raises(oracle.OracleError) as e:
        c.op()
    assert e.value.code == oracle.DECOMPOSITION


def test_parser_quirks(oracle):
    c = oracle.Circuit("V1 1 0 1k\nR1 1 0 1k\n")  # V values take no SI suffix (quirk 7)
    assert c.element(0)[4] == 1.0 and c.element(1)[4] == 1000.0
    c = oracle.Circuit("* t\nv1 in 0 dc 0 PULSE (0 5 1u 100u 100u 5u 10u)\nrin in 0 1.5k\n.control\nfoo\n.endc\n")
    assert c.names == ["I(V1)", "V(in)"] and c.element(1)[4] == 1500.0
    with pytest.raises(oracle.OracleError):
        oracle.Circuit("Q1 1 2 3 mod\n")


def test_trap_closed_form_rc(oracle):
    # TRAP on an RC low-pass driven by a step: v_n = v_{n-1} + (2u - 2v_{n-1}) h/(2RC + h)
    net = "V1 in 0 PULSE(0 1 0 0 0 1 1)\nR1 in out 1k\nC1 out 0 1u\n"
    c = oracle.Circuit(net)
    h = 1e-5
    x, _ = c.tran(h, 1e-3, method=oracle.TRAP)
    vo = x[:, c.index["V(out)"]]
    rc = 1e-3
    v, ref = 0.0, [0.0]
    u0 = 0.0
    for n in range(1, len(vo)):
        v = ((2 * rc / h - 1) * v + u0 + 1.0) / (2 * rc / h + 1)
        u0 = 1.0
        ref.append(v)
    assert np.max(np.abs(vo - ref)) < 1e-14
    xb, _ = c.tran(h, 1e-3, method=oracle.BE)
    assert abs(xb[-1, c.index["V(out)"]] - (1 - (1 / (1 + h / rc)) ** 100)) < 1e-14


def test_adjoint_vs_fd(oracle):
    # discrete adjoint vs central finite differences on an RLC ladder
    net = ("V1 in 0 PULSE(0 1 0 1n 1n 10n 1)\nR1 in m1 20\nL1 m1 n1 2n\nC1 n1 0 0.5p\n"
           "R2 n1 m2 30\nL2 m2 n2 3n\nC2 n2 0 0.3p\n")
    c = oracle.Circuit(net)
    cvec = np.zeros(c.n)
    cvec[c.index["V(n2)"]] = 1.0
    for method in (oracle.TRAP, oracle.BE):
        J, g = c.adjoint(1e-11, 300, cvec, method=method)
        for e in range(c.n_elements):
            t, _, _, _, v = c.element(e)
            if t not in "RCL":
                continue
            d = 1e-6 * v
            c.set_value(e, v + d)
            Jp, _ = c.adjoint(1e-11, 300, cvec, method=method)
            c.set_value(e, v - d)
            Jm, _ = c.adjoint(1e-11, 300, cvec, method=method)
            c.set_value(e, v)
            fd = (Jp - Jm) / (2 * d)
            assert abs(g[e] - fd) <= 1e-6 * abs(fd) + 1e-9 * abs(J) / v, (method, e)


def test_lte_rectifier_accuracy(oracle):
    net = ("V1 in 0 SIN(0 3 1k 0 0 0)\nD1 in out DMOD\nR1 out 0 5k\nC1 out 0 5u\n"
           ".model DMOD D (is=1e-14 n=1)\n")
    c = oracle.Circuit(net)
    cfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    t, x, nt, rej = c.tran_lte(5e-3, 1e-6, 1e-14, 5e-5, 1e-3, 1e-6, cfg=cfg)
    assert abs(t[-1] - 5e-3) < 1e-15 and len(t) < 5000
    xf, _ = c.tran(2e-7, 5e-3, method=oracle.BE, cfg=cfg)  # fine fixed-step reference
    tf = np.arange(xf.shape[0]) * 2e-7
    vo = np.interp(t, tf, xf[:, c.index["V(out)"]])
    assert np.max(np.abs(x[:, c.index["V(out)"]] - vo)) < 2e-2


# ----------------------------------------------- controlled sources (E, G, F, H)
def test_controlled_sources_operating_points(oracle):
    """SPICE definitions of the four linear dependent sources (krets' README
    to-do list, README.md:56-62) on hand-solved circuits."""
    from conftest import CS_NETS
    exp = {"vcvs": {"V(out)": 10.0, "I(E1)": -0.01, "I(V1)": -1e-3},
           "vccs": {"V(out)": -2.0, "I(V1)": -2e-3},
           "cccs": {"I(Vs)": 1e-3, "V(out)": 5.0, "V(mid)": 0.0},
           "ccvs": {"I(Vs)": 1e-3, "V(out)": 2.0, "I(H1)": -2e-3}}
    for name, vals in exp.items():
        c = oracle.Circuit(CS_NETS[name])
        for cost in (0, 1):
            s = sol(c, c.op(oracle.config(cost=cost))[0])
            for k, v in vals.items():
                assert abs(s[k] - v) <= 1e-15 * max(abs(v), 1.0), (name, k, s[k])
    c = oracle.Circuit(CS_NETS["vcvs"])
    assert c.names == ["I(V1)", "V(in)", "I(E1)", "V(out)"]  # branch first, then n+ n- nc+ nc-
    t, p, m, b, v = c.element(2)
    assert (t, p, m, b, v) == ("E", 3, -1, 2, 10.0) and c.element_ctrl(2) == (1, -1, -1)
    c = oracle.Circuit(CS_NETS["cccs"])
    assert c.element(3)[0] == "F" and c.element_ctrl(3) == (-1, -1, c.index["I(Vs)"])


def test_controlled_sources_ac_closed_form(oracle):
    """The active filter of every controlled source against its transfer
    functions: Sallen-Key H(s) = 1 / (s^2 R1 R2 C1 C2 + s C2 (R1 + R2) + 1),
    then I(Vsense) = V(out) / 1k, V(f) = kf I / (1/2k + s Cf),
    V(g) = -gm Rg V(f), V(h) = rh I."""
    from conftest import active_netlist
    r1, r2, c1, c2, kf, gm, rh = 12e3, 27e3, 3.3e-9, 1.5e-9, 3.0, 2e-3, 750.0
    c = oracle.Circuit(active_netlist(r1, r2, c1, c2, 1.0, kf, gm, rh))
    f = np.array([10.0, 1e3, 5e3, 2e4, 1e5, 1e6])
    x = c.ac(f)
    s = 2j * np.pi * f
    vout = 1.0 / (s * s * r1 * r2 * c1 * c2 + s * c2 * (r1 + r2) + 1.0)
    isense = vout / 1e3
    vf = kf * isense / (1.0 / 2e3 + s * 1e-9)
    ref = {"V(out)": vout, "I(Vsense)": isense, "V(f)": vf, "V(g)": -gm * 1e3 * vf, "V(h)": rh * isense,
           "V(in)": np.ones_like(s)}
    for k, v in ref.items():
        got = x[:, c.index[k]]
        assert np.max(np.abs(got - v) / np.abs(v)) < 1e-12, k


def test_controlled_sources_transient_closed_form(oracle):
    """Backward Euler through an E buffer and an H sensing an inductor
    current: v_n = 1 - (1 + h/RC)^-n, i_n = (1/R)(1 - (1 + hR/L)^-n)."""
    net = ("V1 in 0 PULSE(0 1 0 0 0 1 1)\nR1 in a 1k\nC1 a 0 1u\nE1 out 0 a 0 3\nRL out 0 100\n"
           "R2 in b 10\nL1 b 0 1m\nH1 o2 0 L1 5\nR3 o2 0 1k\n")
    c = oracle.Circuit(net)
    h = 1e-5
    x, _ = c.tran(h, 1e-3, method=oracle.BE)
    n = np.arange(x.shape[0])
    v = 1 - (1 + h / 1e-3) ** -n
    i = (1 - (1 + h * 10 / 1e-3) ** -n) / 10
    assert np.max(np.abs(x[:, c.index["V(out)"]] - 3 * v)) < 1e-13
    assert np.max(np.abs(x[:, c.index["I(L1)"]] - i)) < 1e-15
    assert np.max(np.abs(x[:, c.index["V(o2)"]] - 5 * i)) < 1e-14


def test_controlled_sources_in_subcircuits(oracle):
    """An op-amp macro (E) with a finite gain A in a non-inverting amplifier,
    V(out) = A / (1 + A/10), and a current mirror (F) whose sense source is
    the instance's own: controlling nodes map through the pins and a
    controlling element name gets the instance prefix. Ground is passed as a
    pin: krets makes "0" inside a definition a local node."""
    A = 1e5
    net = (".subckt opamp inp inn out gnd\nE1 out gnd inp inn 1e5\n.ends\n"
           ".subckt mirror a b gnd\nVs a gnd 0\nF1 gnd b Vs 2\n.ends\n"
           "V1 in 0 1\nX1 in fb out 0 opamp\nR1 out fb 9k\nR2 fb 0 1k\n"
           "V2 n1 0 1\nR3 n1 ma 1k\nX2 ma mb 0 mirror\nR4 mb 0 1k\n")
    c = oracle.Circuit(net)
    s = sol(c, c.op()[0])
    assert abs(s["V(out)"] - A / (1 + A / 10)) < 1e-12
    assert abs(s["I(V2_s)"] - 1e-3) < 1e-17 and abs(s["V(mb)"] - 2.0) < 1e-13
    assert "I(E1_1)" in c.index


def test_controlled_source_errors(oracle):
    with pytest.raises(oracle.OracleError) as e:
        oracle.Circuit("V1 1 0 1\nR1 1 0 1k\nF1 0 2 Vx 2\nR2 2 0 1k\n")
    assert e.value.code == oracle.ELEMENT_NOT_FOUND
    with pytest.raises(oracle.OracleError) as e:  # R1 has no branch current
        oracle.Circuit("V1 1 0 1\nR1 1 0 1k\nH1 2 0 R1 2\nR2 2 0 1k\n")
    assert e.value.code == oracle.ELEMENT_NOT_FOUND


def test_adjoint_with_controlled_sources_vs_fd(oracle):
    """Discrete adjoint of the active filter: R and C gradients against
    central finite differences (controlled-source gains are constants of the
    step matrix and get no gradient)."""
    from conftest import active_netlist
    c = oracle.Circuit(active_netlist())
    cvec = np.zeros(c.n)
    cvec[c.index["V(g)"]] = 1.0
    cvec[c.index["V(out)"]] = 0.5
    h, N = 1e-6, 120
    for method in (oracle.TRAP, oracle.BE):
        J, g = c.adjoint(h, N, cvec, method=method)
        for e in range(c.n_elements):
            t, _, _, _, v = c.element(e)
            if t not in "RC":
                assert t not in "EGFH" or g[e] == 0.0
                continue
            d = 1e-6 * v
            c.set_value(e, v + d)
            Jp, _ = c.adjoint(h, N, cvec, method=method)
            c.set_value(e, v - d)
            Jm, _ = c.adjoint(h, N, cvec, method=method)
            c.set_value(e, v)
            fd = (Jp - Jm) / (2 * d)
            assert abs(g[e] - fd) <= 1e-6 * abs(fd) + 1e-9 * abs(J) / v, (method, e, g[e], fd)


# ------------------------------------------- single functions of the oracle
# Pulse::value_at known answers: crates/krets-parser/src/elements/voltage_source.rs:369-423
# (krets' tolerance 1e-9 absolute; kept).
PULSE_KAT = [0.0, 5.0, 1e-6, 10e-9, 10e-9, 5e-6, 10e-6]


@pytest.mark.parametrize("t,v", [(0.5e-6, 0.0), (1.005e-6, 2.5), (3.01e-6, 5.0), (6.015e-6, 2.5), (8.0e-6, 0.0),
                                 (11.005e-6, 2.5)])
def test_pulse_known_answers_of_krets(oracle, t, v):
    assert abs(oracle.wave_value(1, PULSE_KAT, t) - v) < 1e-9


# Sinusoidal::value_at known answers: voltage_source.rs:444-496 (offset 1, amplitude 5,
# 1 kHz, delay 1 ms; the last with a 90 degree phase).
@pytest.mark.parametrize("t,phase,v", [(0.5e-3, 0.0, 1.0), (1.25e-3, 0.0, 6.0), (1.5e-3, 0.0, 1.0),
                                       (1.75e-3, 0.0, -4.0), (1e-3, 90.0, 6.0)])
def test_sin_known_answers_of_krets(oracle, t, phase, v):
    assert abs(oracle.wave_value(2, [1.0, 5.0, 1000.0, 1e-3, 0.0, phase], t) - v) < 1e-9


def test_waveform_branches_beyond_krets_tests(oracle):
    """Branches krets' own tests leave open, against hand-computed values."""
    # damping: SIN(0 1 1k 1m 1000 90) one period after the delay: e^-1 * sin(2 pi + pi/2) = e^-1
    assert abs(oracle.wave_value(2, [0.0, 1.0, 1000.0, 1e-3, 1000.0, 90.0], 2e-3) - 0.36787944117144233) < 1e-12
    # a negative phase flips it; the offset adds
    assert abs(oracle.wave_value(2, [0.25, 2.0, 1000.0, 1e-3, 1000.0, -90.0], 2e-3)
               - (0.25 - 2 * 0.36787944117144233)) < 1e-12
    # before the delay the offset holds even with damping and phase
    assert oracle.wave_value(2, [0.25, 2.0, 1000.0, 1e-3, 1000.0, 90.0], 0.999e-3) == 0.25
    # PULSE(1 3 2u 1u 4u 2u 20u): rise, hold, a quarter into the fall, low, and the same points
    # three periods later (fmod wrap); fall from 3 to 1 over 4 us
    p = [1.0, 3.0, 2e-6, 1e-6, 4e-6, 2e-6, 20e-6]
    for k in range(4):
        s = k * 20e-6
        assert abs(oracle.wave_value(1, p, s + 2.5e-6) - 2.0) < 1e-9
        assert oracle.wave_value(1, p, s + 4e-6) == 3.0
        assert abs(oracle.wave_value(1, p, s + 6e-6) - 2.5) < 1e-9
        assert oracle.wave_value(1, p, s + 15e-6) == 1.0
    assert oracle.wave_value(1, p, 1e-6) == 1.0  # before the delay
    # period 0: fmod gives NaN, every comparison fails and the pulse returns v1 (voltage_source.rs:68-93)
    assert oracle.wave_value(1, [1.0, 3.0, 0.0, 1e-6, 1e-6, 1e-6, 0.0], 0.5e-6) == 1.0


# parse_value known answers: crates/krets-parser/src/utils.rs:94-110 (tolerance 1e-15 there)
@pytest.mark.parametrize("tok,v", [("1.5k", 1500.0), ("10u", 10e-6), ("22n", 22e-9), ("1.2p", 1.2e-12),
                                   ("3MEG", 3e6), ("100", 100.0), ("1e-6", 1e-6), ("1.23E-9", 1.23e-9),
                                   ("3meg", 3e6), ("4.7K", 4700.0), ("2f", 2e-15), ("1g", 1e9), ("2T", 2e12),
                                   ("5m", 5e-3)])
def test_parse_value_known_answers_of_krets(oracle, tok, v):
    assert abs(oracle.parse_value(tok) - v) <= 1e-15 * max(1.0, abs(v))


@pytest.mark.parametrize("tok", ["1.5x", "garbage", "", "k", "1.5kk"])
def test_parse_value_errors(oracle, tok):
    with pytest.raises(oracle.OracleError) as e:
        oracle.parse_value(tok)
    assert e.value.code == oracle.PARSE


def test_pnjlim_hand_computed(oracle):
    """SPICE3 pnjlim (Nagel 1975; SPICE3 devsup.c DEVpnjlim), vt = 25 mV, vcrit = 0.6 V."""
    vt, vc = 0.025, 0.6
    v, lim = oracle.pnjlim(1.0, 0.7, vt, vc)  # vold > 0: vold + vt ln(1 + (vnew - vold)/vt) = 0.7 + vt ln 13
    assert lim and abs(v - 0.7641237339365383) < 1e-15
    v, lim = oracle.pnjlim(1.0, -0.1, vt, vc)  # vold <= 0: vt ln(vnew / vt) = vt ln 40
    assert lim and abs(v - 0.09222198635284841) < 1e-15
    v, lim = oracle.pnjlim(0.7, 0.9, vt, vc)  # arg = 1 - 8 <= 0: vcrit
    assert lim and v == 0.6
    assert oracle.pnjlim(0.55, 0.1, vt, vc) == (0.55, False)  # below vcrit
    assert oracle.pnjlim(0.74, 0.7, vt, vc) == (0.74, False)  # within 2 vt of the old value
    assert oracle.pnjlim(-3.0, 0.7, vt, vc) == (-3.0, False)  # reverse bias is not limited


def rc_sine_response(t, tau, w):
    """Closed form of tau x' + x = sin(w t), x(0) = 0 (the RC low-pass driven
    by a sine): x = A sin(w t - phi) + A sin(phi) e^{-t/tau}, A = 1/sqrt(1+(w tau)^2),
    phi = atan(w tau). Returns x, x' and a bound on |x''| near t."""
    A, phi = 1.0 / math.sqrt(1.0 + (w * tau) ** 2), math.atan(w * tau)
    x = A * np.sin(w * t - phi) + A * math.sin(phi) * np.exp(-t / tau)
    x1 = w * A * np.cos(w * t - phi) - A * math.sin(phi) / tau * np.exp(-t / tau)
    x2max = w * w * A + A * math.sin(phi) / tau ** 2 * np.exp(-t / tau)
    return x, x1, x2max


def test_lte_controller_against_closed_form(oracle):
    """Every attempted step of the LTE-controlled backward Euler on an RC
    low-pass driven by a sine, against the closed-form solution.

    The controller estimates the local truncation error h^2/2 |x''| by the
    second divided difference of the last three states. Backward Euler's
    difference quotients (x_n - x_{n-1})/h_n equal the ODE's derivative at
    the step's end, x'(t_n) (up to the global error), so the estimate on this
    circuit is h^2/2 * |2 (x'(t+h) - x'(t)) / (h + h_prev)|, computed here
    from the closed form's x'. It must agree to 5% of the estimate that the
    local amplitude of x'' gives; a factor 2 anywhere in the estimate or the
    tolerance fails that. A step is accepted exactly when its ratio
    (estimate / tolerance) is at most 1; a rejected step retries with
    h * max(1/4, 0.9/sqrt(ratio)), an accepted one proposes
    h * clamp(0.9/sqrt(ratio), 1/4, 2), both capped by hmax and the stop time."""
    tau, f = 1e-4, 1000.0
    w = 2 * math.pi * f
    c = oracle.Circuit("V1 in 0 SIN(0 1 1k 0 0 0)\nR1 in out 1k\nC1 out 0 100n\n")
    reltol, abstol, hmax, tstop = 1e-4, 1e-7, 2e-5, 2e-3
    tr = c.tran_lte_trace(tstop, 1e-8, 1e-14, hmax, reltol, abstol)
    assert (tr[:, 3] == 0).sum() >= 3  # the sine makes the controller reject steps
    t_acc = [0.0]
    worst = 0.0
    for k, (t, h, ratio, outcome, hn) in enumerate(tr):
        assert t == t_acc[-1]
        if ratio >= 0:
            hp = t - t_acc[-2]
            x, x1, x2max = rc_sine_response(np.array([t, t + h]), tau, w)
            tol = reltol * max(abs(x[0]), abs(x[1])) + abstol
            model = 0.5 * h * h * abs(2.0 * (x1[1] - x1[0]) / (h + hp)) / tol
            scale = 0.5 * h * h * x2max[0] / tol
            worst = max(worst, abs(ratio - model) / scale)
            assert (outcome == 1) == (ratio <= 1.0)
            g = 0.9 / math.sqrt(ratio) if ratio > 0 else 2.0
            want = h * max(0.25, g) if outcome == 0 else h * min(2.0, max(0.25, g))
            assert abs(hn - want) <= 1e-15 * want
        else:
            assert outcome == 1 and len(t_acc) == 1  # only the first step has no estimate
            assert hn == h
        if k + 1 < len(tr):
            tn = t + h if outcome == 1 else t
            assert abs(tr[k + 1, 1] - min(hn, hmax, tstop - tn)) <= 1e-12 * hn
        if outcome == 1:
            t_acc.append(t + h)
    assert abs(t_acc[-1] - tstop) < 1e-15
    assert worst < 0.05, worst


def test_transient_source_time_on_a_divider(oracle):
    """krets' backward Euler evaluates the source at the previous step's time
    (transient.rs: the guess map's `time`, quirk 1): a resistor divider's node
    at step n is 3/4 value_at(t_{n-1}); the clean integrators use t_n."""
    h = 0.1e-6
    for wave, kind in (("PULSE(0.5 2.5 1u 1u 2u 3u 10u)", 1), ("SIN(0.5 2 100k 5u 2e4 90)", 2)):
        # waveform parameters take SI suffixes (parse_value: "5u" is 5 * 1e-6, one ulp below 5e-6)
        p = [oracle.parse_value(tok) for tok in wave[wave.index("(") + 1:-1].split()]
        c = oracle.Circuit(f"V1 in 0 0.25 {wave}\nC1 in 0 1n\nR1 in out 1k\nR2 out 0 3k\n")
        k = c.index["V(out)"]
        for method, lag in ((oracle.BE_COMPAT, 1), (oracle.BE, 0), (oracle.TRAP, 0)):
            x, _ = c.tran(h, 400 * h, method=method)
            want = [0.75 * 0.25] + [0.75 * oracle.wave_value(kind, p, (n - lag) * h) for n in range(1, 401)]
            assert np.max(np.abs(x[:, k] - want)) < 1e-14, (wave, method)
