"""CPU checks of the host symbolic pass (no GPU needed).

For every golden netlist, analysis kind and ordering, the exported plan is
checked three ways:
  * its gather lists rebuild exactly krets' MNA matrix and right-hand side, as
    assembled by the oracle (stampable.rs semantics), under the plan's row and
    column permutation;
  * its factor program and level-ordered solves, run in numpy, solve the
    permuted system (compared with numpy.linalg.solve);
  * the transposed solves solve the transposed system.
"""
import math

import numpy as np
import pytest

import plan_emulator as pe
from conftest import CS_NETS, active_netlist, read_circuit

CIRCUITS = [
    "circuit_simulation_farid_n_najm/circuit_simulation_farid_n_najm.cir",
    "voltage_divider/voltage_divider.cir",
    "low_pass_filter/low_pass_filter.cir",
    "high_pass_filter/high_pass_filter.cir",
    "basic_001/basic_001.cir",
    "diode_iv_curve/diode_iv_curve.cir",
    "dual_rc_ladder/dual_rc_ladder.cir",
    "rectifier/rectifier.cir",
    "resistor_ladder_500/resistor_ladder_500.cir",
] + sorted(CS_NETS) + ["active"]


def netlist(case):
    if case == "active":
        return active_netlist()
    return CS_NETS[case] if case in CS_NETS else read_circuit(case)
VT = 1.380649e-23 * 300.0 / 1.602176634e-19


def ladder(sections):
    lines = ["V1 n0 0 1"]
    for k in range(1, sections + 1):
        lines += [f"R{k} n{k-1} m{k} 20", f"L{k} m{k} n{k} 1n", f"C{k} n{k} 0 0.5p"]
    return "\n".join(lines) + "\n"


def values(c, kind, xop=None, f=0.0, h=1e-9, alpha=1.0):
    el, pr = c.elements, c.params

    def diode(e, x):
        p, m = el[e]["plus"], el[e]["minus"]
        vp = 0.5 if (x is None or p < 0) else x[p]
        vm = 0.0 if (x is None or m < 0) else x[m]
        vd = vp - vm
        nvt = pr[e, 1] * VT
        g = pr[e, 0] / nvt * math.exp(vd / nvt)
        return g, pr[e, 0] * (math.exp(vd / nvt) - 1.0) - g * vd

    def y(e):
        t = el[e]["type"]
        if t == 0:
            return 1.0 / pr[e, 0]
        if t == 1:
            return 2j * math.pi * f * pr[e, 0] if kind == pe.AC else alpha * pr[e, 0] / h
        if t == 5:
            return diode(e, xop)[0]
        return 0.0

    def z(e):
        t = el[e]["type"]
        if t == 2:
            return -2j * math.pi * f * pr[e, 0] if kind == pe.AC else -alpha * pr[e, 0] / h
        if t == 0:
            return -pr[e, 0]
        return 1.0

    def contrib(e, k):  # k 5 / 6: +- controlled-source gain
        return [y(e), -y(e), 1.0, -1.0, z(e), pr[e, 0], -pr[e, 0]][k]

    def rhs(e, k):
        t = el[e]["type"]
        if k == 0:
            return pr[e, 1] if (kind == pe.AC and t == 3) else pr[e, 0]
        if k in (1, 2):
            ieq = diode(e, xop)[1]
            return -ieq if k == 1 else ieq
        return 0.0  # companion terms are zero for a zero previous state

    return contrib, rhs


@pytest.mark.parametrize("ordering", [pe.NATURAL, pe.MINDEG, pe.NESTED, pe.BALANCED])
@pytest.mark.parametrize("path", CIRCUITS)
def test_plan_dc_ac_matches_oracle(oracle, path, ordering):
    import krets_amd as k
    text = netlist(path)
    c = k.Circuit(text)
    oc = oracle.Circuit(text)
    assert c.names == oc.names
    # DC at krets' first Newton iteration (empty solution map)
    P = pe.Plan(c.elements, c.n, pe.DC, ordering)
    if P.status:
        with pytest.raises(oracle.OracleError):
            oc.op()
        return
    contrib, rhs = values(c, pe.DC)
    v, M = P.gather(contrib)
    b = P.gather_rhs(rhs)
    A, bref = oc.assemble_dc()
    ro, co = P.arr("row_of", c.n), P.arr("col_of", c.n)
    assert np.allclose(M, A[np.ix_(ro, co)], rtol=1e-15, atol=0)
    assert np.allclose(b, bref[ro], rtol=1e-15, atol=0)
    vf = P.factor(v)
    x = P.solve(vf, b)
    assert np.allclose(x, np.linalg.solve(M, b), rtol=1e-9, atol=1e-14 * np.abs(x).max())
    xt = P.solve_t(vf, b)
    assert np.allclose(xt, np.linalg.solve(M.T, b), rtol=1e-9, atol=1e-14 * np.abs(xt).max())
    # AC at 1234.5 Hz linearised at the operating point
    xop, _ = oc.op()
    PA = pe.Plan(c.elements, c.n, pe.AC, ordering)
    if PA.status:
        with pytest.raises(oracle.OracleError):
            oc.ac([1234.5])
        return
    contrib, rhs = values(c, pe.AC, xop=xop, f=1234.5)
    v, M = PA.gather(contrib)
    b = PA.gather_rhs(rhs)
    A, bref = oc.assemble_ac(xop, 1234.5)
    ro, co = PA.arr("row_of", c.n), PA.arr("col_of", c.n)
    assert np.allclose(M, A[np.ix_(ro, co)], rtol=1e-15, atol=0)
    assert np.allclose(b, bref[ro], rtol=1e-15, atol=0)
    x = PA.solve(PA.factor(v), b)
    assert np.allclose(x, np.linalg.solve(M, b), rtol=1e-9, atol=1e-14 * max(np.abs(x).max(), 1e-300))


@pytest.mark.parametrize("ordering", [pe.MINDEG, pe.NESTED, pe.BALANCED])
@pytest.mark.parametrize("sections", [1, 8, 37])
def test_plan_transient_ladder(oracle, sections, ordering):
    import krets_amd as k
    c = k.Circuit(ladder(sections))
    P = pe.Plan(c.elements, c.n, pe.TRAN, ordering)
    assert P.status == 0
    contrib, _ = values(c, pe.TRAN, h=1e-11, alpha=2.0)
    v, M = P.gather(contrib)
    b = np.random.default_rng(sections).standard_normal(c.n)
    vf = P.factor(v)
    assert np.allclose(P.solve(vf, b), np.linalg.solve(M, b), rtol=1e-9, atol=1e-12)
    assert np.allclose(P.solve_t(vf, b), np.linalg.solve(M.T, b), rtol=1e-9, atol=1e-12)
    if ordering == pe.NESTED and sections == 37:
        # nested dissection keeps the elimination tree shallow on a chain
        assert P.L["nlev"] <= 2 * math.ceil(math.log2(c.n)) + 2
    if ordering == pe.BALANCED and sections == 37:
        # a chain is eliminated from both ends: about half the levels, no fill
        Pm = pe.Plan(c.elements, c.n, pe.TRAN, pe.MINDEG)
        assert P.L["nnz"] == Pm.L["nnz"] and P.L["nlev"] <= Pm.L["nlev"] // 2 + 2


@pytest.mark.parametrize("kind", [pe.DC, pe.TRAN])
@pytest.mark.parametrize("ordering", [pe.MINDEG, pe.NESTED])
def test_plan_rlc_static_pivots(kind, ordering):
    # inductor shorts make the DC matrix's incidence entries cancel exactly
    # under a purely structural pivot order; the representative numeric
    # pivoting must avoid every zero pivot
    import krets_amd as k
    from krets_amd import workloads as W
    b = W.c3_rlc_uneven(6, smin=8, smax=30, seed=7)
    for i in range(6):
        c = k.Circuit(b.netlists[i])
        P = pe.Plan(c.elements, c.n, kind, ordering, c.params, 2e11)
        assert P.status == 0
        contrib, _ = values(c, kind, h=1e-11, alpha=2.0)
        v, M = P.gather(contrib)
        vf = P.factor(v)
        d = vf[P.diag]
        assert np.all(np.isfinite(d)) and np.all(d != 0)
        rhs = np.random.default_rng(i).standard_normal(c.n)
        assert np.allclose(P.solve(vf, rhs), np.linalg.solve(M, rhs), rtol=1e-8, atol=1e-10)


def test_plan_structurally_singular():
    import krets_amd as k
    c = k.Circuit("I1 1 0 1\nR1 1 0 1\n")  # AC current source lacks (ic, ic)
    assert pe.Plan(c.elements, c.n, pe.AC, pe.NESTED).status == 2
    c = k.Circuit("V1 1 0 1\nR1 1 2 1\nC1 2 0 1u G2\n")  # G2 capacitor: empty branch row in DC
    assert pe.Plan(c.elements, c.n, pe.DC, pe.MINDEG).status == 2


LIN_CASES = ["low_pass_filter/transient.cir", "high_pass_filter/transient.cir", "dual_rc_ladder/dual_rc_ladder.cir"]


def lin_bytes(h):
    al = lambda b: (b + 15) & ~15
    return al(2 * h["words_fwd"]) + 8 * (h["nnz"] + 2 * h["n"] + 2 * h["ndyn"] + 5) + 8 * 10 * h["nsrc"] + 64


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("case", LIN_CASES + ["ladder5", "ladder23", "c3", "active"])
def test_lin_plan_step_loop_vs_oracle(oracle, case, method):
    """The compact plan of the LDS-resident transient kernel and its step
    loop, run on the CPU, against the oracle's fixed-step transient."""
    import krets_amd as k
    from krets_amd import workloads as W
    if case.startswith("ladder"):
        text, h, N = ladder(int(case[6:])), 1e-11, 25
    elif case == "c3":
        b = W.c3_rlc_uneven(3, smin=8, smax=40, seed=11)
        text, h, N = b.netlists[2], 10e-12, 25
    elif case == "active":
        text, h, N = active_netlist(), 2e-6, 40
    else:
        text, h, N = read_circuit(case), 50e-6, 25
    c = k.Circuit(text)
    alpha = 2.0 if method == 2 else 1.0
    P = pe.Plan(c.elements, c.n, pe.TRAN, pe.NESTED, c.params, alpha / h)
    lp = pe.LinPlan(c.elements, c.n, c.params, alpha / h)
    assert P.status == 0 and lp.status == 0
    assert lp.h["n"] == c.n and lp.h["nnz"] == P.nnz
    lp.check_indices(len(c.elements), P)
    contrib, _ = values(c, pe.TRAN, h=h, alpha=alpha)
    v, _ = P.gather(contrib)
    vf = P.factor(v).real
    xr, _ = oracle.Circuit(text).tran(h, N * h, method=method)
    x = pe.tran_lin(lp, vf, c.params, xr[0], h, N, method)
    for s in range(N + 1):
        d = np.linalg.norm(xr[s])
        assert np.linalg.norm(x[s] - xr[s]) <= 1e-12 * max(d, 1e-300) + 1e-300, s


def test_lin_plan_sizes():
    """The largest ladder of the uneven transient batch (512 RLC sections)
    fits one CU's 160 KB of LDS in the compact form; diode circuits have no
    compact plan."""
    import krets_amd as k
    c = k.Circuit(ladder(512))
    lp = pe.LinPlan(c.elements, c.n, c.params, 2e11)
    assert lp.status == 0
    lp.check_indices(len(c.elements))
    assert lin_bytes(lp.h) <= 160 * 1024, lin_bytes(lp.h)
    c = k.Circuit(read_circuit("rectifier/rectifier.cir"))
    assert pe.LinPlan(c.elements, c.n, c.params, 1e5).status == 6  # KG_UNEXPECTED: not linear


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("case", ["rlc4", "rc_ladder", "active"])
def test_lin_adjoint_vs_oracle(oracle, case, method):
    """The LDS-resident adjoint's algorithm on the compact plan (transposed
    level solves, element records, checkpoint segments), run on the CPU,
    plus the operating-point term, against the oracle's discrete adjoint."""
    import krets_amd as k
    from krets_amd import workloads as W
    if case == "rlc4":
        text = W.c3_netlist([20.0, 35.0, 50.0, 65.0], [2e-9, 3e-9, 4e-9, 5e-9], [0.3e-12, 0.5e-12, 0.7e-12, 0.9e-12])
        h, N, K, node = 1e-11, 40, 16, "V(n4)"
    elif case == "active":
        text, h, N, K, node = active_netlist(), 2e-6, 40, 8, "V(g)"
    else:
        b = W.c5_rc_ladder(1, sections=12, seed=5)
        text, h, N, K, node = b.netlists[0], 1e-9, 30, 8, "V(n12)"
    c = k.Circuit(text)
    alpha = 2.0 if method == 2 else 1.0
    P = pe.Plan(c.elements, c.n, pe.TRAN, pe.NESTED, c.params, alpha / h)
    lp = pe.LinPlan(c.elements, c.n, c.params, alpha / h)
    lp.check_indices(len(c.elements), P)
    contrib, _ = values(c, pe.TRAN, h=h, alpha=alpha)
    v, _ = P.gather(contrib)
    vf = np.append(P.factor(v).real, 0.0)
    oc = oracle.Circuit(text)
    cvec = np.zeros(c.n)
    cvec[c.names.index(node)] = 1.0
    xr, _ = oc.tran(h, N * h, method=method)
    J, g, xbar0 = pe.adj_lin(lp, vf, c.params, xr[0], h, N, method, cvec, K)
    # operating-point term: lam0 = A_dc^-T xbar0, dJ/dR += (e.lam0)(e.x0) / R^2
    A, _ = oc.assemble_dc()
    lam0 = np.linalg.solve(A.T, xbar0)
    x0 = xr[0]
    for e in range(len(c.elements)):
        el = c.elements[e]
        if el["type"] == 0:
            pl, mi = el["plus"], el["minus"]
            d = (lam0[pl] if pl >= 0 else 0.0) - (lam0[mi] if mi >= 0 else 0.0)
            dx = (x0[pl] if pl >= 0 else 0.0) - (x0[mi] if mi >= 0 else 0.0)
            g[e] += d * dx / c.params[e, 0] ** 2
    Jr, gr = oc.adjoint(h, N, cvec, method=method)
    assert abs(J - Jr) <= 1e-12 * abs(Jr)
    assert np.max(np.abs(g - gr)) <= 1e-9 * np.max(np.abs(gr)), np.max(np.abs(g - gr)) / np.max(np.abs(gr))
