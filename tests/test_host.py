"""CPU checks of the library's host side (no GPU calls): the C ABI exports,
the netlist reader against the oracle's, the krets.toml reader against krets'
own spec files and parser tests, and the parquet writer against the
krets-result schema held by krets' result files.
"""
import json
import math
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, read_circuit

NETLISTS = [
    "circuit_simulation_farid_n_najm/circuit_simulation_farid_n_najm.cir",
    "voltage_divider/voltage_divider.cir",
    "low_pass_filter/low_pass_filter.cir",
    "low_pass_filter/transient.cir",
    "low_pass_filter/ac/ac.cir",
    "high_pass_filter/high_pass_filter.cir",
    "high_pass_filter/transient.cir",
    "basic_001/basic_001.cir",
    "diode_iv_curve/diode_iv_curve.cir",
    "dual_rc_ladder/dual_rc_ladder.cir",
    "rectifier/rectifier.cir",
    "resistor_ladder_500/resistor_ladder_500.cir",
    "resistor_ladder_1000/resistor_ladder_1000.cir",
]


@pytest.fixture(scope="module")
def kg():
    import krets_amd
    krets_amd.lib()
    return krets_amd


def header_functions():
    names = []
    for f in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(kg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_header_function(kg):
    names = header_functions()
    assert len(names) >= 25 and "kg_ac_sweep" in names and "kg_parse_spec" in names
    missing = [n for n in names if not hasattr(kg.lib(), n)]
    assert not missing, missing


# ------------------------------------------------------------- netlists
TYPES = {"R": 0, "C": 1, "L": 2, "V": 3, "I": 4, "D": 5, "E": 6, "G": 7, "F": 8, "H": 9}


def same_circuit(kg, oracle, text):
    c, oc = kg.Circuit(text), oracle.Circuit(text)
    assert c.names == oc.names
    assert c.n_elements == oc.n_elements
    for e in range(c.n_elements):
        t, p, m, b, v = oc.element(e)
        row = c.elements[e]
        assert (int(row["type"]), int(row["plus"]), int(row["minus"]), int(row["branch"])) == (TYPES[t], p, m, b)
        assert (int(row["ctrl_plus"]), int(row["ctrl_minus"]), int(row["ctrl_branch"])) == oc.element_ctrl(e)
        assert c.params[e, 0] == v
    return c


@pytest.mark.parametrize("path", NETLISTS)
def test_netlist_reader_matches_oracle(kg, oracle, path):
    same_circuit(kg, oracle, read_circuit(path))


def test_netlist_reader_quirks(kg, oracle):
    c = same_circuit(kg, oracle, "V1 1 0 1k\nR1 1 0 1k\n")  # V values take no SI suffix
    assert c.params[0, 0] == 1.0 and c.params[1, 0] == 1000.0
    c = same_circuit(kg, oracle, "* t\nv1 in 0 dc 0 PULSE (0 5 1u 100u 100u 5u 10u)\nrin in 0 1.5k\n"
                                 ".control\nfoo\n.endc\n")
    assert c.names == ["I(V1)", "V(in)"] and c.params[1, 0] == 1500.0
    assert c.element_ids == ["V1", "Rin"]  # type letter upper-cased + name (elements.rs:314)
    with pytest.raises(kg.KretsGpuError) as e:
        kg.Circuit("")
    assert e.value.code == 8  # EmptyNetlist
    with pytest.raises(kg.KretsGpuError):
        kg.Circuit("Q1 1 2 3 mod\n")


# parse_value known answers of krets (crates/krets-parser/src/utils.rs:94-110), read through a
# resistor line of the library's netlist reader and the oracle's
SUFFIX_KAT = [("1.5k", 1500.0), ("10u", 10e-6), ("22n", 22e-9), ("1.2p", 1.2e-12), ("3MEG", 3e6), ("100", 100.0),
              ("1e-6", 1e-6), ("1.23E-9", 1.23e-9), ("3meg", 3e6), ("4.7K", 4700.0), ("2f", 2e-15), ("1g", 1e9),
              ("2T", 2e12), ("5m", 5e-3)]


@pytest.mark.parametrize("tok,v", SUFFIX_KAT)
def test_netlist_reader_value_suffixes(kg, oracle, tok, v):
    c = same_circuit(kg, oracle, f"V1 1 0 1\nR1 1 0 {tok}\n")
    assert abs(c.params[1, 0] - v) <= 1e-15 * max(1.0, abs(v))


@pytest.mark.parametrize("tok", ["1.5x", "garbage", "1.5kk"])
def test_netlist_reader_value_suffix_errors(kg, oracle, tok):
    with pytest.raises(kg.KretsGpuError) as e:
        kg.Circuit(f"V1 1 0 1\nR1 1 0 {tok}\n")
    assert e.value.code == 7  # ParseError
    with pytest.raises(oracle.OracleError) as e:
        oracle.Circuit(f"V1 1 0 1\nR1 1 0 {tok}\n")
    assert e.value.code == oracle.PARSE


def test_netlist_reader_controlled_sources(kg, oracle):
    """E, G, F, H lines (SPICE syntax): types, branch currents of E and H,
    controlling nodes and controlling branches, SI-suffixed gains, and the
    index order (branch, n+, n-, nc+, nc-), the same as the oracle's."""
    from conftest import CS_NETS, active_netlist
    for text in list(CS_NETS.values()) + [active_netlist()]:
        same_circuit(kg, oracle, text)
    c = same_circuit(kg, oracle, "E1 o 0 a b 2k\nR1 a 0 1\nR2 b 0 1\nR3 o 0 1\nV1 a b 1\n")
    assert c.names == ["I(E1)", "V(o)", "V(a)", "V(b)", "I(V1)"] and c.params[0, 0] == 2000.0
    el = c.elements[0]
    assert (el["type"], el["branch"], el["ctrl_plus"], el["ctrl_minus"], el["ctrl_branch"]) == (6, 0, 2, 3, -1)
    # controlling element declared after its user; lower-case reference
    c = same_circuit(kg, oracle, "h1 o 0 vsense 10\nR1 o 0 1k\nI1 0 n 1m\nvsense n 0 0\n")
    assert c.element_ids == ["H1", "R1", "I1", "Vsense"]
    assert c.elements[0]["ctrl_branch"] == c.names.index("I(Vsense)")
    # inside subcircuits: the instance's own controlling element
    net = (".subckt mirror a b gnd\nVs a gnd 0\nF1 gnd b Vs 2\nG1 gnd b a gnd 1m\n.ends\n"
           "V2 n1 0 1\nR3 n1 ma 1k\nX2 ma mb 0 mirror\nX3 ma mc 0 mirror\nR4 mb 0 1k\nR5 mc 0 1k\n")
    c = same_circuit(kg, oracle, net)
    f = c.element_ids.index("F2_1")
    assert c.elements[f]["ctrl_branch"] == c.names.index("I(V2_s)")
    assert c.elements[c.element_ids.index("F3_1")]["ctrl_branch"] == c.names.index("I(V3_s)")
    for bad in ("V1 1 0 1\nR1 1 0 1k\nF1 0 2 Vx 2\nR2 2 0 1k\n", "V1 1 0 1\nR1 1 0 1k\nH1 2 0 R1 2\n"):
        with pytest.raises(kg.KretsGpuError) as e:
            kg.Circuit(bad)
        assert e.value.code == 4  # ElementNotFound
    with pytest.raises(kg.KretsGpuError) as e:
        kg.Circuit("E1 o 0 a 2\nR1 o 0 1\n")  # a VCVS needs two controlling nodes
    assert e.value.code == 7


# ------------------------------------------------------------- krets.toml
def test_spec_files_from_krets(kg):
    s = kg.AnalysisSpec.from_file(os.path.join(GOLDEN, "circuits/low_pass_filter/ac/krets.toml"))
    assert s.circuit_path == "ac.cir"
    assert s.analysis == kg.Ac(kg.Linear(10), 0.1, 1000.0)
    s = kg.AnalysisSpec.from_file(os.path.join(GOLDEN, "circuits/high_pass_filter/transient/krets.toml"))
    assert s.circuit_path == "transient.cir" and s.analysis == kg.Transient(10e-6, 2e-3)
    s = kg.AnalysisSpec.from_file(os.path.join(GOLDEN, "circuits/inverter/dc/krets.toml"))
    assert s.circuit_path == "inverter.cir" and s.analysis == kg.Dc("V1", 0.0, 1.0, 0.01)


def test_spec_parser_tests_of_krets(kg):
    # crates/krets-parser/src/analyses.rs:170-259
    a = kg.parse_analysis('\n[transient]\ntime_step = 1e-6\nstop_time = 1e-3\n')
    assert a == kg.Transient(1e-6, 1e-3)
    s = kg.AnalysisSpec.from_str('\ncircuit_path = "any_path/krets.toml"\n\n[analysis.transient]\n'
                                 'time_step = 1e-6\nstop_time = 1e-3\n')
    assert s.circuit_path.endswith("krets.toml") and s.analysis == kg.Transient(1e-6, 1e-3)
    a = kg.parse_analysis('\n[ac]\nsweep = { variation = "Decade", points_per_decade = 10  }\nfstart = 1.0\n'
                          'fstop = 1e6\n')
    assert a == kg.Ac(kg.Decade(10), 1.0, 1e6)
    s = kg.AnalysisSpec.from_str('\ncircuit_path = "any_path/krets.toml"\n\n[analysis.ac]\n'
                                 'sweep = { variation = "Decade", points_per_decade = 5 }\nfstart = 10.0\n'
                                 'fstop = 1000.0\n')
    assert s.analysis == kg.Ac(kg.Decade(5), 10.0, 1000.0)


def test_spec_forms(kg):
    s = kg.AnalysisSpec.from_str('circuit_path = "a.cir"  # comment\nanalysis = "op"\n')
    assert s.analysis == kg.Op()
    s = kg.AnalysisSpec.from_str('circuit_path = "/abs/a.cir"\n[analysis.ac]\n'
                                 'sweep = {variation="Octave",points_per_octave=3}\nfstart = 1\nfstop = 1_000\n')
    assert s.analysis == kg.Ac(kg.Octave(3), 1.0, 1000.0)  # integers are accepted for floats
    s = kg.AnalysisSpec.from_str('circuit_path = "a.cir"\n\n[analysis.dc]\nelement = "I1"\nstart = -1.5\n'
                                 'stop = +2.0\nstep_size = 5e-1\n')
    assert s.analysis == kg.Dc("I1", -1.5, 2.0, 0.5)


@pytest.mark.parametrize("text", [
    "",
    '[analysis.dc]\nelement = "V1"\nstart = 0\nstop = 1\nstep_size = 0.1\n',  # no circuit_path
    'circuit_path = "a"\n',  # no analysis
    'circuit_path = "a"\nanalysis = "tran"\n',
    'circuit_path = "a"\n[analysis.foo]\nx = 1\n',
    'circuit_path = "a"\n[analysis.transient]\ntime_step = 1\n',
    'circuit_path = "a"\n[analysis.transient]\ntime_step = 1\nstop_time = "2"\n',
    'circuit_path = "a"\n[analysis.transient]\ntime_step = 1\nstop_time = 2\n[analysis.dc]\nelement = "V1"\n'
    'start = 0\nstop = 1\nstep_size = 0.1\n',
    'circuit_path = "a"\n[analysis.ac]\nsweep = { variation = "Log", total_points = 3 }\nfstart = 1\nfstop = 2\n',
    'circuit_path = "a"\n[analysis.ac]\nsweep = { variation = "Linear", total_points = 2.5 }\nfstart = 1\n'
    'fstop = 2\n',
    'circuit_path = "a"\n[analysis.ac]\nsweep = { variation = "Linear", total_points = -1 }\nfstart = 1\n'
    'fstop = 2\n',
    'circuit_path = "a"\n[analysis.ac]\nsweep = { variation = "Decade", total_points = 3 }\nfstart = 1\n'
    'fstop = 2\n',
    'circuit_path = "a\n',
    'circuit_path = "a" x\n',
    'circuit_path = "a"\ncircuit_path = "b"\nanalysis = "op"\n',
])
def test_spec_errors(kg, text):
    with pytest.raises(kg.KretsGpuError) as e:
        kg.AnalysisSpec.from_str(text)
    assert e.value.code == 7  # ParseError


def test_circuit_path_resolution(kg, tmp_path):
    from krets_amd.spec import resolve_circuit_path
    d = tmp_path / "case"
    d.mkdir()
    (d / "x.cir").write_text("R1 1 0 1\n")
    spec = d / "krets.toml"
    assert resolve_circuit_path(spec, "x.cir") == str(d / "x.cir")
    assert resolve_circuit_path(spec, str(d / "x.cir")) == str(d / "x.cir")
    with pytest.raises(FileNotFoundError):
        resolve_circuit_path(spec, "missing.cir")


# ------------------------------------------------------------- parquet writer
def read_parquet(path):
    import pyarrow.parquet as pq
    f = pq.ParquetFile(path)
    return f, f.read().to_pydict()


def test_dc_writer_matches_krets_dc_file(tmp_path):
    from krets_amd.results import write_dc_results_to_parquet
    g = json.load(open(os.path.join(GOLDEN, "inverter_dc_schema.json")))
    rng = np.random.default_rng(1)
    keys = [k for k in g["columns"] if k != "step"]
    data = [dict({k: float(rng.normal()) for k in reversed(keys)}, step=float(i)) for i in range(len(g["step"]))]
    path = write_dc_results_to_parquet(data, str(tmp_path / "result"))
    assert path.endswith("result.parquet")
    f, t = read_parquet(path)
    assert f.schema_arrow.names == g["columns"]
    assert [str(x) for x in f.schema_arrow.types] == g["types"]
    assert f.metadata.row_group(0).column(0).compression == g["compression"]
    assert t["step"] == g["step"]
    assert t["V(D)"] == [row["V(D)"] for row in data]


def test_tran_and_op_writers(tmp_path):
    from krets_amd.results import write_op_results_to_parquet, write_tran_results_to_parquet
    data = [{"V(out)": 0.1 * i, "I(V1)": -0.2 * i, "time": 1e-5 * i, "V(in)": 1.0} for i in range(5)]
    data[3].pop("V(in)")
    _, t = read_parquet(write_tran_results_to_parquet(data, str(tmp_path / "tran.parquet")))
    assert list(t) == ["time", "I(V1)", "V(in)", "V(out)"]
    assert t["time"] == [r["time"] for r in data] and t["V(in)"][3] is None
    _, t = read_parquet(write_op_results_to_parquet({"V(b)": 2.0, "I(V1)": -1.0, "V(a)": 1.0},
                                                     str(tmp_path / "op.parquet")))
    assert t == {"I(V1)": [-1.0], "V(a)": [1.0], "V(b)": [2.0]}
    assert write_tran_results_to_parquet([], str(tmp_path / "none.parquet")) is None
    assert not os.path.exists(tmp_path / "none.parquet")


def test_ac_writer_schema_and_polar_form(tmp_path):
    from krets_amd.results import write_ac_results_to_parquet
    data = [{"frequency": complex(10.0, 0.0), "V(b)": 3 + 4j, "I(V1)": -1 + 0j, "V(a)": -1j},
            {"frequency": complex(20.0, 0.0), "V(b)": 1 + 1j, "I(V1)": 0j, "V(a)": -2 - 2j}]
    _, t = read_parquet(write_ac_results_to_parquet(data, str(tmp_path / "ac.parquet")))
    assert list(t) == ["frequency", "I(V1)_mag", "I(V1)_phase_deg", "V(a)_mag", "V(a)_phase_deg", "V(b)_mag",
                       "V(b)_phase_deg"]
    assert t["frequency"] == [10.0, 20.0]
    assert t["V(b)_mag"] == [5.0, math.sqrt(2.0)]
    assert t["V(b)_phase_deg"][0] == pytest.approx(math.degrees(math.atan(4.0 / 3.0)), abs=1e-13)
    assert t["V(b)_phase_deg"][1] == 45.0 and t["I(V1)_phase_deg"] == [180.0, 0.0]
    assert t["V(a)_phase_deg"] == [-90.0, -135.0] and t["V(a)_mag"] == [1.0, math.sqrt(8.0)]


def test_ac_writer_on_golden_columns(tmp_path):
    """Rebuild the complex values of krets' AC golden file from its own
    magnitude and phase, write them, and read back the same columns (the
    polar form round-trips to a few ulps)."""
    from krets_amd.results import write_ac_results_to_parquet
    g = json.load(open(os.path.join(GOLDEN, "low_pass_ac_result.json")))
    cols = {k: [float(v) for v in vals] for k, vals in g["columns"].items()}
    names = ["I(V1)", "V(in)", "V(out)"]
    data = []
    for i, f in enumerate(cols["frequency"]):
        row = {"frequency": complex(f, 0.0)}
        for n in names:
            m, p = cols[n + "_mag"][i], math.radians(cols[n + "_phase_deg"][i])
            row[n] = complex(m * math.cos(p), m * math.sin(p))
        data.append(row)
    _, t = read_parquet(write_ac_results_to_parquet(data, str(tmp_path / "result.parquet")))
    assert list(t) == list(g["columns"])
    assert t["frequency"] == cols["frequency"]
    for k in t:
        np.testing.assert_allclose(t[k], cols[k], rtol=1e-14, atol=1e-12)


# ------------------------------------------------------------- subcircuits
NESTED = """
xdiv1 10 7 0 vdivide

.subckt vdivide 1 2 3
xr1 1 2 myresistor
xr2 2 3 myresistor
.ends


.subckt myresistor plus minus
r0 plus minus 1k
.ends
"""


def test_subcircuit_fixture_of_krets(kg, oracle):
    """circuits/voltage_divider/subckt/voltage_divider.cir: the instance card
    precedes the definition; pins 1 2 3 map to 10 7 0."""
    text = read_circuit("voltage_divider/subckt/voltage_divider.cir")
    c = same_circuit(kg, oracle, text)
    assert c.names == ["V(10)", "V(7)", "I(V1)", "V(1)"]
    assert c.element_ids == ["Rdiv1_1", "Rdiv1_2", "V1"]
    assert list(c.params[:, 0]) == [10e3, 5e3, 10.0]
    x, _ = oracle.Circuit(text).op()
    assert list(x) == [0.0, 0.0, 0.0, 10.0]  # the divider has no source; V1 drives node 1 alone


def test_subcircuit_nested_of_krets(kg, oracle):
    # crates/krets-parser/tests/test_parse_netlist.rs:78-102: two elements
    c = same_circuit(kg, oracle, NESTED)
    assert c.element_ids == ["Rdiv1_r1_0", "Rdiv1_r2_0"] and c.names == ["V(10)", "V(7)"]


def test_subcircuit_expansion_equals_flat_netlist(kg, oracle):
    """An RC ladder built from nested instances is the netlist written out by
    hand with krets' naming: pins map to the instance's nodes, other nodes
    and element names take "<instance>_" as prefix, in definition order."""
    sub = ("V1 in 0 DC 1 AC 1\nXa in m sec2\nXb m out sec2\nRL out 0 2k\n"
           ".subckt sec2 a b\nXs1 a mid sec\nXs2 mid b sec\n.ends\n"
           ".subckt sec p q\nR1 p q 1k\nC1 q 0 1n\n.ends\n")
    flat = ("V1 in 0 DC 1 AC 1\n"
            "Ra_s1_1 in a_mid 1k\nCa_s1_1 a_mid a_s1_0 1n\nRa_s2_1 a_mid m 1k\nCa_s2_1 m a_s2_0 1n\n"
            "Rb_s1_1 m b_mid 1k\nCb_s1_1 b_mid b_s1_0 1n\nRb_s2_1 b_mid out 1k\nCb_s2_1 out b_s2_0 1n\n"
            "RL out 0 2k\n")
    cs, cf = same_circuit(kg, oracle, sub), same_circuit(kg, oracle, flat)
    assert cs.names == cf.names and cs.element_ids == cf.element_ids
    assert np.array_equal(cs.elements, cf.elements) and np.array_equal(cs.params, cf.params)
    # ground inside a definition is a local node unless it is a pin: "a_s1_0"
    assert "V(a_s1_0)" in cs.names and "V(0)" not in cs.names


@pytest.mark.parametrize("text", [
    "V1 a 0 1\nX1 a b nodef\n",                                    # undefined definition
    "V1 a 0 1\nX1 a sec\n.subckt sec p q\nR1 p q 1\n.ends\n",        # pin count
    "V1 a 0 1\nX1 a b loop\n.subckt loop p q\nX2 p q loop\n.ends\n",  # recursive
    "V1 a 0 1\n.subckt bad p\n.model D1 D\n.ends\n",                 # not an element
])
def test_subcircuit_errors(kg, oracle, text):
    with pytest.raises(kg.KretsGpuError) as e:
        kg.Circuit(text)
    assert e.value.code == 7  # ParseError
    with pytest.raises(oracle.OracleError):
        oracle.Circuit(text)


def test_header_is_c99_and_host_entry_points_work_from_c(kg, tmp_path):
    """include/krets_gpu.h compiles as strict C99 (gcc -std=c99 -pedantic
    -Wall -Wextra -Werror), and a C caller of the host-only entry points
    (netlist reader, element tables, plan export, default config, spec
    reader) gets what the Python binding gets."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "abi_check")
    lib_dir = os.path.join(ROOT, "krets_amd")
    r = subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "host_c99", "abi_check.c"), "-o",
                        exe, "-L" + lib_dir, "-Wl,-rpath," + lib_dir, "-l:libkrets_gpu.so"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    cir = os.path.join(GOLDEN, "circuits", "low_pass_filter", "ac", "ac.cir")
    toml = os.path.join(GOLDEN, "circuits", "low_pass_filter", "ac", "krets.toml")
    r = subprocess.run([exe, cir, toml], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    c = kg.Circuit(open(cir).read())
    assert lines[0] == "parse 0" and lines[1] == f"size {c.n} {c.n_elements}"
    assert [l.split()[2] for l in lines if l.startswith("unknown ")] == c.names
    assert [l.split()[2] for l in lines if l.startswith("element ")] == c.element_ids
    rows = [l.split()[2:] for l in lines if l.startswith("row ")]
    for e, row in enumerate(rows):
        el = c.elements[e]
        assert [int(v) for v in row[:4]] == [int(el["type"]), int(el["plus"]), int(el["minus"]), int(el["branch"])]
        assert float(row[4]) == c.params[e, 0]
    plan = next(l for l in lines if l.startswith("plan ")).split()
    assert plan[1] == "1" and int(plan[2]) == c.n
    assert next(l for l in lines if l.startswith("config ")) == "config 0.001 9.9999999999999998e-13 9.9999999999999995e-07 300 1"
    spec = next(l for l in lines if l.startswith("spec ")).split()
    assert spec[1] == "0" and spec[2] == "2" and spec[3].endswith(".cir")


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N above 1 exits with status 2 and says why,
    before it touches torch or a GPU."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""
