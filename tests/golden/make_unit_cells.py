#!/usr/bin/env python3
"""Writes tests/golden/unit_cells.json: the data of LPM-CPP's UnitCell
(include/unit_cell.h:44-122) -- el_mapping, nneighbors, dim, neighbor_cutoff
and particle_volume of the five lattices -- evaluated from the reference
header's own expressions.

The header is read as text (nothing of the reference is compiled, imported or
run): each set*() body's assignments are parsed and evaluated with C's
arithmetic (integer literals stay integers, int / int truncates, sqrt and pow
are the C library's, left-to-right association), so every value is the double
the reference's constructor computes.  Values are stored as hexadecimal floats
so the comparison is bit for bit.

Run in the container that holds the reference:
    python3 tests/golden/make_unit_cells.py [/path/to/LPM-CPP]
"""
import ast
import json
import math
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCE = "include/unit_cell.h"
# set*() function -> lattice id of include/lpm_gpu.h (LPMG_SQUARE2D ... LPMG_BCC3D)
FUNCS = {"setSquare": 0, "setHexagon": 1, "setSimpleCubic": 2, "setFCC": 3, "setBCC": 4}
RADII = [0.05, 0.1, 0.12, 0.13, 0.5, 1.0]


class CDouble:
    """A C arithmetic value: int or double, with C's promotion rules."""

    def __init__(self, v, is_int):
        self.v, self.is_int = v, is_int


def _binop(op, a, b):
    both_int = a.is_int and b.is_int
    x, y = (a.v, b.v) if both_int else (float(a.v), float(b.v))
    if isinstance(op, ast.Add):
        r = x + y
    elif isinstance(op, ast.Sub):
        r = x - y
    elif isinstance(op, ast.Mult):
        r = x * y
    elif isinstance(op, ast.Div):
        if both_int:  # C truncates toward zero
            q = abs(x) // abs(y)
            r = q if (x >= 0) == (y > 0) else -q
        else:
            r = x / y
    else:
        raise ValueError(f"operator {op}")
    return CDouble(r, both_int)


def c_eval(expr, radius):
    """Evaluate a C expression of the header (numbers, radius, sqrt, pow)."""
    expr = expr.replace("std::", "")
    # C float literals such as "1." or "2.0" are Python floats; "12" stays int
    tree = ast.parse(expr, mode="eval")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant):
            return CDouble(n.value, isinstance(n.value, int))
        if isinstance(n, ast.Name):
            if n.id == "radius":
                return CDouble(radius, False)
            raise ValueError(f"name {n.id}")
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, ast.USub):
            a = ev(n.operand)
            return CDouble(-a.v, a.is_int)
        if isinstance(n, ast.BinOp):
            return _binop(n.op, ev(n.left), ev(n.right))
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name):
            args = [float(ev(a).v) for a in n.args]
            if n.func.id == "sqrt":
                return CDouble(math.sqrt(*args), False)
            if n.func.id == "pow":
                return CDouble(math.pow(*args), False)
        raise ValueError(f"unsupported expression {ast.dump(n)}")

    return ev(tree)


def parse(text):
    out = {}
    for fn, lat in FUNCS.items():
        m = re.search(r"void\s+UnitCell::" + fn + r"\s*\(\s*\)\s*\{(.*?)\n\}", text, re.S)
        if not m:
            raise SystemExit(f"{fn} not found in {SOURCE}")
        body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
        stmts = {}
        for s in body.split(";"):
            s = " ".join(s.split())
            if "=" in s:
                k, v = s.split("=", 1)
                stmts[k.strip()] = v.strip()
        el = [e.strip() for e in stmts["el_mapping"].strip("{}").split(",")]
        cut = [e.strip() for e in stmts["neighbor_cutoff"].strip("{}").split(",")]
        assert len(el) == 9 and len(cut) == 2, fn
        rec = dict(function=fn, dim=int(stmts["dim"]), nneighbors=int(stmts["nneighbors"]),
                   el_mapping=[float(c_eval(e, 1.0).v).hex() for e in el],
                   el_mapping_expr=el, cutoff_expr=cut, volume_expr=stmts["particle_volume"], radii={})
        for r in RADII:
            rec["radii"][repr(r)] = dict(cutoff=[float(c_eval(e, r).v).hex() for e in cut],
                                         volume=float(c_eval(stmts["particle_volume"], r).v).hex())
        out[str(lat)] = rec
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, SOURCE)) as f:
        data = parse(f.read())
    doc = dict(source=f"LPM-CPP {SOURCE} (UnitCell::setSquare ... setBCC)",
               note="values as float.hex(); lattice keys are LPMG_SQUARE2D..LPMG_BCC3D", lattices=data)
    path = os.path.join(HERE, "unit_cells.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
