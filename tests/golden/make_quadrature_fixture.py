"""Writes quadrature_rules.json: the points and weights of the reference's
Gauss-Legendre rules (fem/QuadratureRule.cc), 2 to 20 points, in the order
the reference lists them.  The rule tables are data; this script reads them
from the reference source as text and keeps only the numbers.

    python tests/golden/make_quadrature_fixture.py /root/reference
"""
import json
import os
import re
import sys


def main(ref):
    src = open(os.path.join(ref, "fem", "QuadratureRule.cc")).read()
    rules = {}
    for m in re.finditer(r'compare\("(\d+)pt"\)(.*?)(?=compare\("\d+pt"\)|\Z)', src, re.S):
        n = int(m.group(1))
        body = m.group(2)
        arrays = {}
        for name in ("points", "weights"):
            a = re.search(name + r"\[\d*\]\s*=\s*\{(.*?)\}", body, re.S)
            arrays[name] = [float(v) for v in a.group(1).replace("\n", " ").split(",")]
            assert len(arrays[name]) == n, (n, name)
        rules[str(n)] = arrays
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "quadrature_rules.json")
    with open(out, "w") as f:
        json.dump({"source": "fem/QuadratureRule.cc", "rules": rules}, f, indent=1)
    print(f"wrote {len(rules)} rules to {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
