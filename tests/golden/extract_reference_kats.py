"""Write tests/golden/reference_kats.json from the reference test-suite.

Reads (as text, with ``ast``; nothing is imported or executed) the literal
known-answer data of
  tests/test_gmg.py:8-170        coarse DOF index arrays (create_coarse_inds)
  tests/test_compliance.py:8-113 end-to-end compliance goldens (filters,
                                 optimizers, heat)
  tests/test_linsolver.py:10-28  the 3x3 linear-system known answer
  tests/test_passive.py:7-10     the passive-element (el_flags) golden
and stores it as JSON.  Run once in a container that has the reference at
/root/reference; the JSON is committed, so the tests never need the reference.
"""
from __future__ import annotations

import ast
import json
from pathlib import Path

REF = Path("/root/reference/tests")
OUT = Path(__file__).resolve().parent / "reference_kats.json"


def _lit(node):
    """Literal value of an AST node, allowing array([...]) and None names."""
    if isinstance(node, ast.Call) and getattr(node.func, "id", "") == "array":
        return _lit(node.args[0])
    if isinstance(node, ast.Name):
        return {"None": None}.get(node.id, node.id)
    if isinstance(node, ast.Tuple):
        return [_lit(e) for e in node.elts]
    if isinstance(node, ast.List):
        return [_lit(e) for e in node.elts]
    return ast.literal_eval(node)


def _parametrize_rows(path: Path, func: str):
    tree = ast.parse(path.read_text())
    for fn in ast.walk(tree):
        if isinstance(fn, ast.FunctionDef) and fn.name == func:
            for dec in fn.decorator_list:
                if isinstance(dec, ast.Call) and getattr(dec.func, "attr", "") == "parametrize":
                    names = [s.strip() for s in ast.literal_eval(dec.args[0]).split(",")]
                    rows = [_lit(r) for r in dec.args[1].elts]
                    return names, rows
    raise KeyError(func)


def main():
    out = {}
    names, rows = _parametrize_rows(REF / "test_gmg.py", "test_coarse_inds")
    out["coarse_inds"] = {"names": names, "rows": rows,
                          "source": "tests/test_gmg.py:8-83"}
    names, rows = _parametrize_rows(REF / "test_compliance.py", "test_compliance_filters")
    out["compliance_filters"] = {"names": names, "rows": rows,
                                 "source": "tests/test_compliance.py:7-35 (assert to 5 decimals)"}
    names, rows = _parametrize_rows(REF / "test_compliance.py", "test_compliance_optimizers")
    out["compliance_optimizers"] = {
        "names": names, "rows": rows,
        "case": "nelx 60, nely 20, volfrac 0.5, ft 1, rmin 2.4, matrix filter, mbb_2d",
        "source": "tests/test_compliance.py:62-86 (assert to 5 decimals)"}
    names, rows = _parametrize_rows(REF / "test_compliance.py", "test_compliance_heat")
    out["compliance_heat"] = {"names": names, "rows": rows,
                              "source": "tests/test_compliance.py:88-113 (3 decimals)"}
    names, rows = _parametrize_rows(REF / "test_passive.py", "test_compliance_1")
    out["compliance_passive"] = {
        "names": names, "rows": rows,
        "passive": "circle of element flags 1: center (nelx/3 - 1, nely/2 - 1), radius nely/3, "
                   "element (i, j) = divmod(e, nely) inside when (i-cx)^2 + (j-cy)^2 <= r^2",
        "source": "tests/test_passive.py:7-31 (assert to 2 decimals)"}
    out["linsystem_3x3"] = {"A": [[2, 1, 0], [1, 2, 0], [0, 0, 1]], "b": [1, 2, 0],
                            "x": [0.0, 1.0, 0.0], "source": "tests/test_linsolver.py:20-26"}
    OUT.write_text(json.dumps(out, indent=1))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
