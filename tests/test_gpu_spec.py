"""End to end on the GPU: krets.toml spec -> netlist -> HIP solve ->
result.parquet, as krets-cli runs it (crates/krets-cli/src/main.rs:33-126).

The AC case reproduces krets' own result file
circuits/low_pass_filter/ac/result.parquet (kept as JSON): same columns in
the same order, bit-equal frequencies, magnitudes to 1e-12 relative and phases
to 1e-9 degrees. The other analyses are checked against the oracle to 1e-12
(linear OP/DC) and 1e-9 (transient) normwise per row.
"""
import json
import os
import shutil
import sys

import numpy as np
import pytest

from conftest import GOLDEN, read_circuit

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def nerr(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    d = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (d if d > 0 else 1.0)


@pytest.fixture(scope="module")
def kg():
    import krets_amd
    return krets_amd


def read_parquet(path):
    import pyarrow.parquet as pq
    return pq.read_table(path).to_pydict()


def copy_case(tmp_path, rel_dir):
    dst = tmp_path / os.path.basename(rel_dir)
    shutil.copytree(os.path.join(GOLDEN, "circuits", rel_dir), dst)
    return dst


def test_spec_ac_reproduces_krets_result_file(kg, tmp_path):
    say("spec end to end: low_pass_filter/ac")
    d = copy_case(tmp_path, "low_pass_filter/ac")
    result, path = kg.run(str(d / "krets.toml"))
    assert path == str(d / "result.parquet") and len(result) == 10
    t = read_parquet(path)
    g = json.load(open(os.path.join(GOLDEN, "low_pass_ac_result.json")))
    cols = {k: np.array([float(v) for v in vals]) for k, vals in g["columns"].items()}
    assert list(t) == list(g["columns"])
    assert np.array_equal(t["frequency"], cols["frequency"])
    for k in cols:
        if k.endswith("_mag"):
            assert np.max(np.abs(np.array(t[k]) / cols[k] - 1)) < 1e-12, k
        elif k.endswith("_phase_deg"):
            assert np.max(np.abs(np.array(t[k]) - cols[k])) < 1e-9, k


def test_spec_transient_high_pass(kg, oracle, tmp_path):
    say("spec end to end: high_pass_filter/transient")
    d = copy_case(tmp_path, "high_pass_filter/transient")
    result, path = kg.run(str(d / "krets.toml"))
    t = read_parquet(path)
    oc = oracle.Circuit(read_circuit("high_pass_filter/transient/transient.cir"))
    xr, _ = oc.tran(10e-6, 2e-3)
    assert list(t) == ["time"] + sorted(oc.names)
    assert len(t["time"]) == xr.shape[0] == 201
    assert t["time"] == [i * 10e-6 for i in range(201)]
    x = np.array([t[k] for k in oc.names]).T
    assert max(nerr(x[i], xr[i]) for i in range(len(x))) < 1e-9


def test_spec_op_and_dc(kg, oracle, tmp_path):
    say("spec end to end: voltage divider OP and DC sweep")
    text = read_circuit("voltage_divider/voltage_divider.cir")
    (tmp_path / "div.cir").write_text(text)
    oc = oracle.Circuit(text)
    (tmp_path / "krets.toml").write_text('circuit_path = "div.cir"\nanalysis = "op"\n')
    _, path = kg.run(str(tmp_path / "krets.toml"))
    t = read_parquet(path)
    xr, _ = oc.op()
    assert list(t) == sorted(oc.names)
    assert nerr([t[k][0] for k in oc.names], xr) < 1e-12
    (tmp_path / "krets.toml").write_text('circuit_path = "div.cir"\n\n[analysis.dc]\nelement = "V1"\n'
                                         'start = 0.0\nstop = 1.0\nstep_size = 0.1\n')
    _, path = kg.run(str(tmp_path / "krets.toml"))
    t = read_parquet(path)
    xr = oc.dc("V1", 0.0, 1.0, 0.1)
    assert list(t) == sorted(oc.names + ["step"])
    assert t["step"] == [float(i) for i in range(11)]
    x = np.array([t[k] for k in oc.names]).T
    assert max(nerr(x[i], xr[i]) for i in range(11)) < 1e-12


SUBCKT_LADDER = ("V1 in 0 PULSE(0 1 0 1n 1n 10n 1) AC 1\nXa in m 0 sec2\nXb m out 0 sec2\nRL out 0 2k\n"
                 ".subckt sec2 a b g\nXs1 a mid g sec\nXs2 mid b g sec\n.ends\n"
                 ".subckt sec p q g\nR1 p q 1k\nC1 q g 1n\nL1 q x 1u\nR2 x g 10k\n.ends\n")


def test_subcircuit_netlist_on_gpu(kg, oracle):
    """A netlist built from nested subcircuit instances (ground passed as a
    pin): OP, AC and transient on the GPU against the oracle."""
    say("subcircuit netlist: OP, AC, transient")
    c = kg.Circuit(SUBCKT_LADDER)
    oc = oracle.Circuit(SUBCKT_LADDER)
    assert c.names == oc.names and len(c.element_ids) == 18
    s = kg.Solver(c)
    r = s.solve(kg.Op())
    xr, _ = oc.op()
    assert nerr([r[k] for k in oc.names], xr) < 1e-12
    ac = kg.Ac(kg.Linear(3), 10.0, 1e6)
    f = ac.generate_frequencies()
    res = s.solve(ac)
    xa = oc.ac(f)
    for j in range(3):
        assert nerr([res[j][k] for k in oc.names], xa[j]) < 1e-12
    res = s.solve(kg.Transient(1e-9, 100e-9))
    xt, _ = oc.tran(1e-9, 100e-9)
    assert max(nerr([res[i][k] for k in oc.names], xt[i]) for i in range(len(xt))) < 1e-9
