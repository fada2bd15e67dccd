"""Loads tests/golden/cases.json into frames and canonical expected results."""
import json
import math
from pathlib import Path

import numpy as np

import oracle as orc

CASES = json.loads((Path(__file__).resolve().parent / "golden" / "cases.json").read_text())


def _dec(x):
    if x is None:
        return None
    if isinstance(x, str):
        return float("nan") if x == "nan" else float.fromhex(x)
    return x


def frame_of(case):
    frame = {}
    for name, c in case["columns"].items():
        tag = c["tag"]
        vals = [_dec(v) for v in c["values"]]
        valid = np.array([v is not None for v in vals], dtype=bool)
        fill = False if tag == "Bool" else 0
        arr = np.array([fill if v is None else v for v in vals], dtype=orc.NP_DTYPE[tag])
        frame[name] = (tag, arr, None if valid.all() else valid)
    return frame


def expected_of(case, frame):
    """Expected result in oracle.groupby's form (first-occurrence order)."""
    exp = case["expected"]
    out = {}
    keys = [k[0] for k in case["plan"]["keys"]]
    for name in keys:
        tag = case["columns"][name]["tag"]
        vals = [_dec(v) for v in exp[name]]
        valid = np.array([v is not None for v in vals], dtype=bool)
        fill = False if tag == "Bool" else 0
        out[name] = (tag, np.array([fill if v is None else v for v in vals], dtype=orc.NP_DTYPE[tag]), valid)
    for a in case["plan"]["aggs"]:
        alias = a["output_alias"]
        op = a["op"]
        vals = [_dec(v) for v in exp[alias]]
        valid = np.array([v is not None for v in vals], dtype=bool)
        if op in ("Len", "Count"):
            tag = "U32"
        else:
            src = (case["expr_tags"][alias] if a.get("kind") == "Expression"
                   else case["columns"][a["input_col"]]["tag"])
            tag = "F64" if op == "Mean" and src != "F32" else src
        arr = np.array([0 if v is None else v for v in vals], dtype=orc.NP_DTYPE[tag])
        out[alias] = (tag, arr, valid)
    out["__n_groups__"] = case["n_groups"]
    return out
