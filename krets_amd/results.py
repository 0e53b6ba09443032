"""Parquet result files in krets-result's schema
(crates/krets-result/src/lib.rs:9-207), written with pyarrow.

- OP: one row, one float64 column per key. krets writes them in HashMap
  order, which is not stable; here they are sorted.
- DC: the sorted keys (unknowns and "step") as float64 columns.
- Transient: as DC with "time" moved first.
- AC: "frequency" (the real part of the "frequency" entry), then for every
  other key in sorted order `<key>_mag` = sqrt(re*re + im*im) and
  `<key>_phase_deg` = atan2(im, re) * 180 / pi.
A key missing from a row is a null. Files are ZSTD-compressed, as polars'
ParquetWriter writes by default.
"""
import math
import os

import numpy as np


def ensure_parquet_extension(filename: str) -> str:
    """lib.rs:9-17."""
    return filename if os.path.splitext(filename)[1] == ".parquet" else f"{filename}.parquet"


def _write(columns, filename):
    import pyarrow as pa
    import pyarrow.parquet as pq
    filename = ensure_parquet_extension(os.fspath(filename))
    table = pa.table({k: pa.array(v, type=pa.float64()) for k, v in columns})
    pq.write_table(table, filename, compression="zstd")
    return filename


def _real(v):
    return None if v is None else float(v)


def write_op_results_to_parquet(data, filename):
    """lib.rs:19-51."""
    return _write([(k, [_real(data[k])]) for k in sorted(data)], filename)


def _headers(data):
    return sorted({k for row in data for k in row})


def write_dc_results_to_parquet(data, filename):
    """lib.rs:53-88. Returns None for an empty result (no file), as krets."""
    if not data:
        return None
    return _write([(h, [_real(row.get(h)) for row in data]) for h in _headers(data)], filename)


def write_tran_results_to_parquet(data, filename):
    """lib.rs:90-129."""
    if not data:
        return None
    hs = _headers(data)
    if "time" in hs:
        hs.remove("time")
        hs.insert(0, "time")
    return _write([(h, [_real(row.get(h)) for row in data]) for h in hs], filename)


def _mag_phase(v):
    re, im = float(v.real), float(v.imag)
    return math.sqrt(re * re + im * im), math.atan2(im, re) * 180.0 / math.pi


def write_ac_results_to_parquet(data, filename):
    """lib.rs:131-207."""
    if not data:
        return None
    cols = [("frequency", [None if row.get("frequency") is None else float(np.real(row["frequency"]))
                           for row in data])]
    for h in (h for h in _headers(data) if h != "frequency"):
        mp = [None if row.get(h) is None else _mag_phase(row[h]) for row in data]
        cols.append((f"{h}_mag", [None if x is None else x[0] for x in mp]))
        cols.append((f"{h}_phase_deg", [None if x is None else x[1] for x in mp]))
    return _write(cols, filename)


def write_results_to_parquet(analysis, result, filename):
    """The writer krets-cli picks for each AnalysisResult variant
    (crates/krets-cli/src/main.rs:101-126)."""
    from .solver import Ac, Dc, Op, Transient
    if isinstance(analysis, Op):
        return write_op_results_to_parquet(result, filename)
    if isinstance(analysis, Dc):
        return write_dc_results_to_parquet(result, filename)
    if isinstance(analysis, Ac):
        return write_ac_results_to_parquet(result, filename)
    if isinstance(analysis, Transient):
        return write_tran_results_to_parquet(result, filename)
    raise TypeError(f"unknown analysis {analysis!r}")
