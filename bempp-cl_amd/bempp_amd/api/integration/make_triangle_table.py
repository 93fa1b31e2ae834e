"""Write triangle_rules.npz: the symmetric triangle rules of orders 1-20 as numeric data.

The reference serves its regular quadrature from the HyENA tables of symmetric Gauss rules
(bempp_cl/api/integration/triangle_gauss.py:36-173 for the point counts and table offsets,
175-2225 barycentric coordinates, 2226-2925 weights, 2928-2940 `rule`). Results equal to the
reference's need the same points in the same order, so the rules are kept as data: this script
reads the reference file as text (Python's ast, nothing of it is imported or run), applies the
reference's indexing (rule of order o: n = points_per_order[o - 1] points from table row
points_address[n - 1]; reference point = (2nd, 3rd) barycentric coordinate; weight = table
weight / 2) and stores, per order, points (2, n) and weights (n,).

Usage (in a checkout that has the reference sources):
    python make_triangle_table.py [path/to/triangle_gauss.py]
"""

import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "triangle_rules.npz")
DEFAULT_SOURCE = "/root/reference/bempp_cl/api/integration/triangle_gauss.py"


def read_tables(path):
    """The module-level arrays of the reference file, as Python lists (literal evaluation only)."""
    tree = ast.parse(open(path).read())
    tables = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and isinstance(node.targets[0], ast.Name):
            value = node.value
            if isinstance(value, ast.Call) and value.args:
                tables[node.targets[0].id] = ast.literal_eval(value.args[0])
    return tables


def build(path=DEFAULT_SOURCE):
    t = read_tables(path)
    per_order = np.array(t["points_per_order"], dtype=np.int64)
    address = np.array(t["points_address"], dtype=np.int64)
    coords = np.array(t["coords"], dtype=np.float64)
    weights = np.array(t["weights"], dtype=np.float64)
    data = {"points_per_order": per_order}
    for order in range(1, len(per_order) + 1):
        n = int(per_order[order - 1])
        row = int(address[n - 1])
        if row < 0:
            raise ValueError(f"the table has no {n}-point rule for order {order}")
        bary = coords[3 * row:3 * (row + n)].reshape((n, 3))
        data[f"p{order}"] = np.ascontiguousarray(np.vstack([bary[:, 1], bary[:, 2]]))
        data[f"w{order}"] = 0.5 * weights[row:row + n]
    np.savez(OUT, **data)
    return OUT


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SOURCE))
