#!/usr/bin/env python3
"""Collect the golden values that tandem's own unit tests hold into reference_golden.json.

Sources (read as text; nothing from the reference is executed):
  test/basis.cpp     : warp-and-blend nodes, triangle and tetrahedron, P1..P4 ("Triangle nodes P*",
                       "Tet nodes P*" subcases)
  test/quadrules.cpp : Gauss-Jacobi(5,1,3) vs Maple, collapsed triangle rule (order 3) vs Maple,
                       autorule sizes
  test/form.cpp      : diagonal of the modal tetrahedron mass matrix (degrees 0..3)

Usage: python3 extract_golden.py [/path/to/tandem]
"""
import json
import os
import re
import sys

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def numbers(text):
    return [float(x) for x in re.findall(NUM, text)]


def block(text, start_marker, end_marker):
    i = text.index(start_marker)
    j = text.index(end_marker, i)
    return text[i:j]


def node_block(text, title, dim):
    b = block(text, 'SUBCASE("%s")' % title, "checkNodes(")
    b = b[b.index("refNodes") + len("refNodes"):]
    vals = []
    for m in re.finditer(r"\{([^{}]*)\}", b):
        inner = m.group(1)
        # allow "1.0 / 3.0" style entries
        parts = [p.strip() for p in inner.split(",") if p.strip()]
        if len(parts) != dim:
            continue
        vals.append([float(eval(p, {"__builtins__": {}}, {})) for p in parts])
    return vals


def main(root):
    out = {}
    basis = open(os.path.join(root, "test/basis.cpp")).read()
    for dim, name in ((2, "Triangle"), (3, "Tet")):
        for p in (1, 2, 3, 4):
            out["wab_%dd_p%d" % (dim, p)] = node_block(basis, "%s nodes P%d" % (name, p), dim)
    quad = open(os.path.join(root, "test/quadrules.cpp")).read()
    gj = block(quad, 'SUBCASE("Gauss Jacobi vs Maple")', "SUBCASE(\"Triangle vs Maple\")")
    vals = [float(x) for x in re.findall(r"Approx\((" + NUM + r")\)", gj)]
    out["gauss_jacobi_5_1_3"] = {"points": vals[:5], "weights": vals[5:10]}
    tri = block(quad, 'SUBCASE("Triangle vs Maple")', "auto testOrder")
    lines = re.findall(r"rule\.(points|weights)\(\)\[(\d)\](?:\[(\d)\])? == doctest::Approx\((" + NUM + r")\)", tri)
    pts = [[None, None] for _ in range(4)]
    wts = [None] * 4
    for kind, i, j, v in lines:
        if kind == "points":
            pts[int(i)][int(j)] = float(v)
        else:
            wts[int(i)] = float(v)
    out["tensor_triangle_order3"] = {"points": pts, "weights": wts}
    auto = re.findall(r"simplexQuadratureRule<(\d)u>\((\d+)\)\.size\(\) == (\d+)", quad)
    out["autorule_sizes"] = [[int(d), int(o), int(n)] for d, o, n in auto]
    form = open(os.path.join(root, "test/form.cpp")).read()
    mm = block(form, "refModalMassMatrix = {", "};")
    out["modal_mass_tet_diag"] = numbers(mm.split("{", 1)[1])
    return out


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = main(root)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_golden.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", path, {k: len(v) for k, v in data.items()})
