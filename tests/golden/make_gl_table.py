"""Writes gl_table.json: the Gauss-Legendre nodes and weights the reference tabulates
(src/utils/Quadrule/Tabulated.cpp, Quadrule::GLTable) for n = 1..10, 12, 14, 16, 32, as the decimal
strings printed there.  Reads the reference source as text; run from the repository root:
    python tests/golden/make_gl_table.py /root/reference"""
import json
import re
import sys


def parse(text):
    body = text[text.index("void Quadrule::GLTable"):]
    body = body[:body.index("default:")]
    rules = {}
    for m in re.finditer(r"case\s+(\d+):(.*?)\}\s*return;", body, re.S):
        n = int(m.group(1))
        nodes, weights = [None] * n, [None] * n
        for stmt in m.group(2).split(";"):
            lits = re.findall(r"(nodes|weights)\s*\[\s*(\d+)\s*\]", stmt)
            val = re.findall(r"=\s*\(?\s*(-?\d+\.\d+)\s*\)?\s*$", stmt.strip())
            if not lits or not val:
                continue
            v = val[0]
            # "a[i] = -( a[j] = v )" sets a[j] = v and a[i] = -v; "a[i] = a[j] = v" sets both to v
            neg = "-(" in stmt.replace(" ", "")
            for k, (arr, idx) in enumerate(lits):
                target = nodes if arr == "nodes" else weights
                s = v
                if neg and k == 0:
                    s = v[1:] if v.startswith("-") else "-" + v
                target[int(idx)] = s
        if None in nodes or None in weights:
            raise ValueError(f"rule {n} incomplete")
        rules[n] = {"nodes": nodes, "weights": weights}
    return rules


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(f"{root}/src/utils/Quadrule/Tabulated.cpp").read()
    rules = {n: r for n, r in parse(text).items() if n <= 32}
    with open("tests/golden/gl_table.json", "w") as f:
        json.dump({"source": "src/utils/Quadrule/Tabulated.cpp (Quadrule::GLTable)", "rules": rules}, f, indent=0)
    print("wrote", sorted(rules))
