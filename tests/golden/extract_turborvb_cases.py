"""Write turborvb_cases.json: the golden inputs and outputs held by the reference's TurboRVB
comparison tests (tests/test_comparison_with_turborvb_AE.py, tests/test_comparison_with_turborvb_ECP.py).

The test sources are parsed with ``ast`` (nothing is imported or executed).  For every test
function it records the numeric literals bound to names (coordinates, golden values, Dt,
geminal matrices), the in-place coordinate edits ``new_r_*_carts[i] = [...]`` and the fixture
file names it opens.  The pickled Jastrow fixtures are *not* loaded; the tests map each
``jastrow_data_<suffix>.pkl`` to the TurboRVB text file ``turborvb_WF_<suffix>.txt`` it was
generated from.

Run in a container that has the reference checkout:  python tests/golden/extract_turborvb_cases.py
"""

import ast
import json
import os
import sys

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def _literal(node):
    """Numeric literal, list literal, np.array(list) or np.array(list).T, else None."""
    transpose = False
    if isinstance(node, ast.Attribute) and node.attr == "T":
        node, transpose = node.value, True
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute) and node.func.attr in ("array", "asarray"):
        if node.args:
            node = node.args[0]
    try:
        val = ast.literal_eval(node)
    except Exception:
        return None
    if isinstance(val, (int, float)) or (isinstance(val, list)):
        if transpose:
            import numpy as np

            val = np.asarray(val, dtype=float).T.tolist()
        return val
    return None


def _cases(path):
    tree = ast.parse(open(path).read())
    out = {}
    for fn in tree.body:
        if not isinstance(fn, ast.FunctionDef) or not fn.name.startswith("test_"):
            continue
        case = {"values": {}, "edits": [], "files": []}
        for node in ast.walk(fn):
            if isinstance(node, ast.Constant) and isinstance(node.value, str):
                if node.value.endswith((".h5", ".pkl")):
                    case["files"].append(node.value)
        for stmt in fn.body:
            if isinstance(stmt, ast.Assign) and len(stmt.targets) == 1:
                t = stmt.targets[0]
                val = _literal(stmt.value)
                if isinstance(t, ast.Name) and val is not None:
                    case["values"][t.id] = val
                elif isinstance(t, ast.Subscript) and isinstance(t.value, ast.Name) and val is not None:
                    idx = ast.literal_eval(t.slice)
                    case["edits"].append([t.value.id, idx, val])
        out[fn.name] = case
    return out


def main():
    cases = {}
    for name in ("test_comparison_with_turborvb_AE.py", "test_comparison_with_turborvb_ECP.py"):
        for k, v in _cases(os.path.join(REF, name)).items():
            v["source"] = name
            cases[k] = v
    with open(os.path.join(HERE, "turborvb_cases.json"), "w") as f:
        json.dump(cases, f, indent=1, sort_keys=True)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    sys.exit(main())
