#!/usr/bin/env python3
"""Extract the published simplex quadrature tables used by tandem into a plain text data file.

The tables are published data:
  * F.D. Witherden, P.E. Vincent, "On the identification of symmetric quadrature rules for
    finite element methods", Comput. Math. Appl. 69 (2015) 1232-1241 (triangle, tetrahedron;
    points on the [-1,1]^D reference simplex, weights summing to its volume).
  * J. Jaskowiec, N. Sukumar, "High-order cubature rules for tetrahedra", Int. J. Numer.
    Methods Eng. 121 (2020) 2418-2436 (points on the unit tetrahedron, weights summing to 1).

tandem ships them as generated C++ initialiser lists (src/quadrules/WitherdenVincent2015_Tri.cpp,
WitherdenVincent2015_Tet.cpp, JaskowiecSukumar2020.cpp).  This script reads those files as text,
keeps the orders this project needs and writes them, unchanged in value and point order, to
``simplex_rules.txt`` next to this script.  The conversion to the unit reference simplex
(TabulatedRules.h:66-104) is done by the code that reads the file.

Usage: python3 extract_rules.py /path/to/tandem/src/quadrules
"""
import os
import re
import sys

# (file, family tag, dimension, orders to keep)
SOURCES = [
    ("WitherdenVincent2015_Tri.cpp", "WV", 2, range(1, 21)),
    ("WitherdenVincent2015_Tet.cpp", "WV", 3, range(1, 11)),
    ("JaskowiecSukumar2020.cpp", "JS", 3, range(1, 14)),
]

_TOKEN = re.compile(r"[{}]|[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def parse_table(text):
    """Parse the brace-nested initialiser of a std::map<order, pair<points, weights>>."""
    start = text.index("{", text.index(">>")) if ">>" in text else text.index("{")
    body = text[start:]
    stack = [[]]
    for tok in _TOKEN.findall(body):
        if tok == "{":
            stack.append([])
        elif tok == "}":
            done = stack.pop()
            stack[-1].append(done)
            if len(stack) == 1:
                break
        else:
            stack[-1].append(tok)
    table = stack[0][0]
    rules = {}
    for entry in table:
        order = int(entry[0])
        points, weights = entry[1]
        rules[order] = (points, weights)
    return rules


def main(srcdir, out):
    lines = [
        "# Simplex quadrature rules (published data, see extract_rules.py for sources).",
        "# rule <family> <dim> <order> <npoints>, then one line per point: coordinates weight.",
        "# WV: [-1,1]^D reference simplex coordinates; JS: unit tetrahedron coordinates.",
    ]
    for fname, fam, dim, orders in SOURCES:
        with open(os.path.join(srcdir, fname)) as f:
            text = f.read()
        # skip the declaration line (std::map<unsigned, std::pair<...>>) before the table
        text = text[text.index(fam == "JS" and "JaskowiecSukumar2020{" or
                               ("WitherdenVincent2015_Tri{" if dim == 2 else "WitherdenVincent2015_Tet{")):]
        text = text[text.index("{"):]
        rules = parse_table(" >> " + text)
        for order in orders:
            if order not in rules:
                continue
            pts, wts = rules[order]
            assert len(pts) == len(wts)
            lines.append("rule %s %d %d %d" % (fam, dim, order, len(pts)))
            for p, w in zip(pts, wts):
                assert len(p) == dim
                lines.append(" ".join(p) + " " + w)
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/quadrules"
    main(src, os.path.join(os.path.dirname(os.path.abspath(__file__)), "simplex_rules.txt"))
