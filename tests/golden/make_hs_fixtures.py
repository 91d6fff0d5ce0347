"""Extract the golden Hubbard-Stratonovich fields of the reference tests as data.

Reads ``test/hs_ref.py`` of the reference checkout *as text* (it is parsed with
``ast``, never imported or executed) and writes the integer arrays
``hs_init[geometry]`` / ``hs_end[geometry]`` to ``hs_ref.json`` beside this
script.  Source: test/hs_ref.py:5-48 (hs_init["square"]), :49 (triangular reuses
the square initial field), :51 (honeycomb), :1417 (kagome), :3423-3466
(hs_end["square"]), :3467-3511 (hs_end["triangular"]), :3512 (honeycomb),
:4878 (kagome).  The run that produced hs_end is
test/test_gen_1band_hub_unified.py:119-148: create_batch(seed=1234,
n_sweep_warm=20, n_sweep_meas=20) with every other generator default.

Usage: python tests/golden/make_hs_fixtures.py [/path/to/reference]
"""
import ast
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
GEOMETRIES = ("square", "triangular", "honeycomb", "kagome")


def _array_literal(node):
    # np.array([...]) -> the list literal
    assert isinstance(node, ast.Call) and node.args, ast.dump(node)[:80]
    return ast.literal_eval(node.args[0])


def main():
    src = open(os.path.join(REF, "test", "hs_ref.py")).read()
    tree = ast.parse(src)
    found = {}
    for stmt in tree.body:
        if not isinstance(stmt, ast.Assign) or len(stmt.targets) != 1:
            continue
        tgt = stmt.targets[0]
        if not isinstance(tgt, ast.Subscript):
            continue
        name = tgt.value.id
        key = ast.literal_eval(tgt.slice)
        if isinstance(stmt.value, ast.Call):
            found[(name, key)] = _array_literal(stmt.value)
        elif isinstance(stmt.value, ast.Subscript):  # alias, e.g. triangular = square
            found[(name, key)] = found[(stmt.value.value.id, ast.literal_eval(stmt.value.slice))]
    out = {
        "source": "reference test/hs_ref.py (parsed as text)",
        "run": {"geometry": "square", "Nx": 4, "Ny": 4, "U": 6.0, "mu": 0.0, "tp": 0.0,
                "dt": 0.1, "L": 40, "n_matmul": 8, "n_delay": 16, "period_eqlt": 8,
                "n_sweep_warm": 20, "n_sweep_meas": 20, "seed": 1234},
        "hs_init": {g: found[("hs_init", g)] for g in GEOMETRIES},
        "hs_end": {g: found[("hs_end", g)] for g in GEOMETRIES},
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hs_ref.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path)


if __name__ == "__main__":
    main()
