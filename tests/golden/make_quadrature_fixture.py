"""Write tests/golden/quadratures.json from the reference's table literals.

Reads ``superspinsim/quadratures.py`` of the reference as text and parses
the numeric list literals with ``ast.literal_eval``; nothing from the
reference is imported or executed.  Run once in a container that has the
reference checkout:  python tests/golden/make_quadrature_fixture.py
"""

import ast
import json
import os
import re

SRC = "/root/reference/superspinsim/quadratures.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "quadratures.json")


def main():
    text = open(SRC).read()
    out = {"samples": {}, "weights": {}}
    for kind in ("samples", "weights"):
        pattern = kind + r'\["([^"]+)"\] = np\.array\((\[.*?\]),\s*dtype'
        for key, body in re.findall(pattern, text, flags=re.S):
            out[kind][key] = ast.literal_eval(body)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(OUT, {k: sorted(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
