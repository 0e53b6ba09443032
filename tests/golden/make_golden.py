"""Regenerate the golden fixtures under tests/golden/ from the krets repository.

Run once in a container that holds the reference checkout (default path
/root/reference, or $KRETS_REFERENCE). It copies the small netlists that
krets' own solver tests load (crates/krets-solver/tests/*.rs) and converts the
one full-precision golden vector, circuits/low_pass_filter/ac/result.parquet,
to JSON with every float written in round-trip form, plus the column layout
of circuits/inverter/dc/result.parquet (a MOSFET circuit, so only its schema
and step column serve as a DC-writer fixture). Nothing here runs krets.
"""
import json
import os
import shutil
import sys

REF = os.environ.get("KRETS_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

NETLISTS = [
    "circuit_simulation_farid_n_najm/circuit_simulation_farid_n_najm.cir",
    "voltage_divider/voltage_divider.cir",
    "voltage_divider/subckt/voltage_divider.cir",
    "low_pass_filter/low_pass_filter.cir",
    "low_pass_filter/transient.cir",
    "low_pass_filter/ac/ac.cir",
    "low_pass_filter/ac/krets.toml",
    "high_pass_filter/high_pass_filter.cir",
    "high_pass_filter/transient.cir",
    "high_pass_filter/transient/krets.toml",
    "high_pass_filter/transient/transient.cir",
    "inverter/dc/krets.toml",
    "basic_001/basic_001.cir",
    "diode_iv_curve/diode_iv_curve.cir",
    "dual_rc_ladder/dual_rc_ladder.cir",
    "rectifier/rectifier.cir",
    "resistor_ladder_500/resistor_ladder_500.cir",
    "resistor_ladder_1000/resistor_ladder_1000.cir",
]


def main():
    if not os.path.isdir(REF):
        sys.exit(f"reference checkout not found at {REF}")
    out = os.path.join(HERE, "circuits")
    for rel in NETLISTS:
        dst = os.path.join(out, rel)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copyfile(os.path.join(REF, "circuits", rel), dst)
    import pyarrow.parquet as pq
    t = pq.read_table(os.path.join(REF, "circuits/low_pass_filter/ac/result.parquet")).to_pydict()
    with open(os.path.join(HERE, "low_pass_ac_result.json"), "w") as f:
        json.dump({"source": "circuits/low_pass_filter/ac/result.parquet",
                   "analysis": {"variation": "Linear", "total_points": 10, "fstart": 0.1,
                                "fstop": 1000.0},
                   "columns": {k: [repr(float(v)) for v in vals] for k, vals in t.items()}},
                  f, indent=1)
    pf = pq.ParquetFile(os.path.join(REF, "circuits/inverter/dc/result.parquet"))
    t = pf.read().to_pydict()
    with open(os.path.join(HERE, "inverter_dc_schema.json"), "w") as f:
        json.dump({"source": "circuits/inverter/dc/result.parquet",
                   "columns": pf.schema_arrow.names,
                   "types": [str(x) for x in pf.schema_arrow.types],
                   "compression": pf.metadata.row_group(0).column(0).compression,
                   "step": t["step"]}, f, indent=1)


if __name__ == "__main__":
    main()
