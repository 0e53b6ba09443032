"""Opt-in diode parasitics: series resistance and junction charge (CPU).

krets reads a diode model's is, rs and n and uses is and n only
(crates/krets-parser/src/models/diode.rs:14-37); it has no junction charge.
kg_parse_netlist_ext with KG_NETLIST_DIODE_PARASITICS (and the oracle's
ref_parse_ext with flag 1) adds both: rs as an internal anode node V(<id>:a)
and a resistor <id>:rs numbered after krets' own unknowns and elements, and
the SPICE2 junction charge (cjo, vj, m, fc, tt; Nagel 1975) in the clean
transient integrators and the AC small-signal stamp.

The oracle's restatement is pinned here against what the mathematics fixes:
the operating point against the circuit equations with the series resistor,
the AC response against the closed form H = 1 / (1 + R (G_d + j w C(v))) of
an RC divider with the junction as C, on both sides of the depletion
formula's switch at fc vj, and the trapezoidal transient against an
independent ODE solve (scipy Radau) of C(v) dv/dt = (v_in - v) / R - I_d(v),
converging at second order (BE at first order).
"""
import numpy as np
import pytest

VT = 1.380649e-23 * 300.0 / 1.602176634e-19  # k T / q at 300 K, as krets' diode model
IS, N, CJO, VJ, M, FC, TT = 1e-14, 1.5, 2e-12, 0.7, 0.4, 0.5, 5e-9
MODEL = f".model dmod D(is={IS} n={N} rs=10 cjo={CJO} vj={VJ} m={M} fc={FC} tt={TT})\n"


def capacitance(v):
    """SPICE2 junction capacitance: depletion CJO (1 - v/VJ)^-M below FC VJ,
    its linear continuation above, plus the diffusion term TT dI/dv."""
    gd = IS / (N * VT) * np.exp(v / (N * VT))
    if v < FC * VJ:
        cj = CJO * (1.0 - v / VJ) ** (-M)
    else:
        cj = CJO / (1.0 - FC) ** (1.0 + M) * (1.0 - FC * (1.0 + M) + M * v / VJ)
    return cj + TT * gd, gd


def test_flag_off_reads_what_krets_reads(oracle):
    import krets_amd as kg
    net = "V1 in 0 5\nR1 in a 1k\nD1 a 0 dmod\n" + MODEL
    for c in (kg.Circuit(net), kg.Circuit(net, diode_parasitics=False)):
        assert c.names == ["I(V1)", "V(in)", "V(a)"] and c.element_ids == ["V1", "R1", "D1"]
        assert (c.params[2, 2:] == 0).all()  # rs and the junction-charge parameters are not read
    assert oracle.Circuit(net).names == c.names


def test_flag_on_numbering_and_tables(oracle):
    import krets_amd as kg
    from krets_amd._lib import D, R
    net = "V1 in 0 5\nR1 in a 1k\nD1 a 0 dmod\nD2 0 a dmod\n" + MODEL
    c = kg.Circuit(net, diode_parasitics=True)
    oc = oracle.Circuit(net, diode_parasitics=True)
    # krets' unknowns and elements keep their numbers; each diode's internal node and resistor follow
    assert c.names == oc.names == ["I(V1)", "V(in)", "V(a)", "V(D1:a)", "V(D2:a)"]
    assert c.element_ids == ["V1", "R1", "D1", "D2", "D1:rs", "D2:rs"]
    ia, d1a, d2a = 2, 3, 4
    e = c.elements
    assert (e["type"][2], e["plus"][2], e["minus"][2]) == (D, d1a, -1)
    assert (e["type"][3], e["plus"][3], e["minus"][3]) == (D, d2a, ia)
    assert (e["type"][4], e["plus"][4], e["minus"][4]) == (R, ia, d1a)
    assert (e["type"][5], e["plus"][5], e["minus"][5]) == (R, -1, d2a)  # D2's anode is ground
    assert c.params[4, 0] == c.params[5, 0] == 10.0
    assert list(c.params[2, :7]) == [IS, N, CJO, VJ, M, FC, TT]
    for k in range(6):
        t, p, m, b, v = oc.element(k)
        assert (p, m) == (e["plus"][k], e["minus"][k]) and t == "RCLVIDEGFH"[e["type"][k]]
    # a model without rs adds no node
    c2 = kg.Circuit("V1 in 0 5\nR1 in a 1k\nD1 a 0 dm\n.model dm D(is=1e-14 cjo=1p)\n", diode_parasitics=True)
    assert c2.names == ["I(V1)", "V(in)", "V(a)"] and c2.params[2, 2] == 1e-12


def test_series_resistance_operating_point(oracle):
    """V1 - R1 - (Rs + junction): the solution satisfies the circuit's own
    equations, I = (V - V_a) / R1 = (V_a - V_j) / Rs = IS (exp(V_j / (N VT)) - 1),
    with Newton run to tight tolerances (krets' 1e-3 stops it ~1e-4 short)."""
    net = "V1 in 0 5\nR1 in a 1k\nD1 a 0 dmod\n" + MODEL
    oc = oracle.Circuit(net, diode_parasitics=True)
    x, _ = oc.op(oracle.config(reltol=1e-14, vabstol=1e-15, iabstol=1e-18))
    va, vj = x[oc.index["V(a)"]], x[oc.index["V(D1:a)"]]
    i1 = (5.0 - va) / 1e3
    assert abs(-x[oc.index["I(V1)"]] - i1) <= 1e-12 * i1
    assert abs((va - vj) / 10.0 - i1) <= 1e-9 * i1
    assert abs(IS * np.expm1(vj / (N * VT)) - i1) <= 1e-9 * i1
    # without the flag the junction sits directly on V(a): a different operating point
    x0, _ = oracle.Circuit(net).op()
    assert x0[2] < va - 0.03


@pytest.mark.parametrize("vdc", [-2.0, -0.5, 0.0, 0.3, 0.34, 0.36, 0.5, 0.6])
def test_junction_capacitance_ac_closed_form(oracle, vdc):
    """Biased diode behind R: V(a) / V(in) = 1 / (1 + R (G_d + j w C(v_a)))
    at the operating point v_a, below and above the switch at FC VJ = 0.35 V."""
    r = 1e3
    net = f"V1 in 0 DC {vdc} AC 1\nR1 in a {r}\nD1 a 0 dm\n.model dm D(is={IS} n={N} cjo={CJO} vj={VJ} m={M} fc={FC} tt={TT})\n"
    oc = oracle.Circuit(net, diode_parasitics=True)
    x, _ = oc.op()
    f = np.array([1e3, 1e6, 1e8, 1e9])
    h = oc.ac(f)[:, oc.index["V(a)"]]
    c, gd = capacitance(x[oc.index["V(a)"]])
    ref = 1.0 / (1.0 + r * (gd + 2j * np.pi * f * c))
    assert np.max(np.abs(h / ref - 1.0)) < 1e-12
    assert np.abs(h[-1] - 1.0) > 0.1  # the capacitance does act at 1 GHz


def test_capacitance_is_continuous_at_the_switch():
    """The linear continuation above FC VJ meets the depletion formula with
    equal value and slope (what makes the charge C^1)."""
    v0, d = FC * VJ, 1e-7
    lo, hi = capacitance(v0 - d)[0], capacitance(v0 + d)[0]
    assert abs(hi - lo) <= 1e-6 * lo
    s_lo = (capacitance(v0 - d)[0] - capacitance(v0 - 2 * d)[0]) / d
    s_hi = (capacitance(v0 + 2 * d)[0] - capacitance(v0 + d)[0]) / d
    assert abs(s_hi - s_lo) <= 1e-3 * abs(s_lo)


def junction_ode(a, f, r, t_end, ts):
    from scipy.integrate import solve_ivp

    def rhs(t, y):
        v = y[0]
        return [((a * np.sin(2 * np.pi * f * t) - v) / r - IS * np.expm1(v / (N * VT))) / capacitance(v)[0]]

    sol = solve_ivp(rhs, (0.0, t_end), [0.0], method="Radau", rtol=1e-11, atol=1e-15, t_eval=ts,
                    max_step=ts[1] - ts[0])
    assert sol.success
    return sol.y[0]


def test_junction_charge_transient_against_ode(oracle):
    """SIN-driven R + junction: the clean integrators converge to the ODE
    solution, TRAP at second order (error ratio 4 per halving of h), BE at
    first order (ratio 2)."""
    a, f, r, t_end = 0.9, 20e6, 1e3, 100e-9
    net = f"V1 in 0 SIN(0 {a} {f} 0 0 0)\nR1 in a {r}\nD1 a 0 dm\n.model dm D(is={IS} n={N} cjo={CJO} vj={VJ} m={M} fc={FC} tt={TT})\n"
    oc = oracle.Circuit(net, diode_parasitics=True)
    ia = oc.index["V(a)"]
    errs = {}
    for method in (oracle.TRAP, oracle.BE):
        errs[method] = []
        for h in (0.1e-9, 0.05e-9, 0.025e-9):
            x, _ = oc.tran(h, t_end, method=method)
            ts = np.minimum(np.arange(x.shape[0]) * h, t_end)
            errs[method].append(np.max(np.abs(x[:, ia] - junction_ode(a, f, r, t_end, ts))))
    e = errs[oracle.TRAP]
    assert e[-1] < 2e-6 and all(3.9 < e[k] / e[k + 1] < 4.1 for k in range(2)), e
    e = errs[oracle.BE]
    assert all(1.8 < e[k] / e[k + 1] < 2.2 for k in range(2)), e
    # krets' own integrator (BE_COMPAT) has no junction charge: it ignores the parameters
    xc, _ = oc.tran(0.1e-9, t_end, method=oracle.BE_COMPAT)
    xn, _ = oracle.Circuit(net).tran(0.1e-9, t_end, method=oracle.BE_COMPAT)
    assert np.array_equal(xc, xn)


def test_grid_integrator_reproduces_the_lte_and_fixed_step_runs(oracle):
    """ref_tran_grid (clean BE over a given grid, used to check the library's
    LTE runs on their own accepted grids) reproduces the oracle's LTE run on
    its grid and the fixed-step BE run on a uniform one."""
    net = ("V1 in 0 SIN(0 2 20e6 0 0 0)\nR1 in a 100\nD1 a out dm\nRL out 0 5k\nCL out 0 5p\n"
           ".model dm D(is=1e-14 n=1.5 rs=20 cjo=2p vj=0.7 m=0.4 fc=0.5 tt=5n)\n")
    oc = oracle.Circuit(net, diode_parasitics=True)
    cfg = oracle.config(vabstol=1e-9, iabstol=1e-12)
    tr, xr, _, _ = oc.tran_lte(100e-9, 0.01e-9, 1e-16, 1e-9, 1e-3, 1e-6, cfg=cfg)
    xg, _ = oc.tran_grid(tr, cfg=cfg)
    assert np.max(np.abs(xg - xr)) <= 1e-13 * np.max(np.abs(xr))
    xb, _ = oc.tran(0.1e-9, 50e-9, method=oracle.BE, cfg=cfg)
    xu, _ = oc.tran_grid(np.arange(501) * 0.1e-9, cfg=cfg)
    assert np.max(np.abs(xu - xb)) <= 1e-13 * np.max(np.abs(xb))
