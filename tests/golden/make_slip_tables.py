"""Extract the crystal-frame slip systems (directions b and plane normals n,
as integer Miller-index triples) of the reference's slip types 1 (fcc),
2 (bcc), 6 (roters), 7 (bcc12) and 8 (bcc48) from its constant assignments
(src/mod_crystals.f:438-1207) into slip_tables.json, the fixture the
slip-table tests compare the oracle's and the library's tables with.

Each assignment reads `c_array(num)%bi(s,k)=<value>` with <value> one of
the named scale factors of the block (f, f2, f3, f112, f211, f123, f213,
f312, z0 or 0, possibly negated) or a literal such as 0.d0; a factor is turned back into its integer
multiple of the block's unit 1/sqrt(m) (m = |index|^2), so the fixture holds
small integers only.  Run in the container that holds the reference:
    python tests/golden/make_slip_tables.py
"""
import json
import math
import os
import re

SRC = "/root/reference/src/mod_crystals.f"
HERE = os.path.dirname(os.path.abspath(__file__))

ASSIGN = re.compile(r"c_array\(num\)%(bi|ni)\(\s*(\d+)\s*,\s*(\d)\s*\)\s*=\s*(-?)\s*([A-Za-z0-9_.]+)")
FACTOR = re.compile(r"^\s*(f\w*|z0)\s*=\s*([0-9.dD]+)\s*/\s*d?sqrt\(\s*([0-9.dD]+)\s*\)")


def fnum(s):
    return float(s.lower().replace("d", "e"))


def block(lines, first, last):
    """the tables assigned in source lines first..last (1-based)"""
    fac = {"0": 0.0, "z0": 0.0}
    tab = {"bi": {}, "ni": {}}
    for ln in lines[first - 1:last]:
        if ln[:1] in "cC*!":
            continue
        m = FACTOR.match(ln)
        if m:
            fac[m.group(1)] = fnum(m.group(2)) / math.sqrt(fnum(m.group(3)))
            continue
        m = ASSIGN.search(ln)
        if m:
            arr, s, k, neg, name = m.groups()
            base = fac[name] if name in fac else fnum(name)  # a named factor or a literal
            v = base * (-1.0 if neg else 1.0)
            tab[arr].setdefault(int(s), [0.0, 0.0, 0.0])[int(k) - 1] = v
    ns = max(tab["bi"])
    out = {}
    for arr in ("bi", "ni"):
        rows = [tab[arr][s] for s in range(1, ns + 1)]
        ints = []
        for r in rows:
            # smallest integer triple proportional to r: scale by the
            # largest unit among 1/sqrt(2), 1/sqrt(3), 1/sqrt(6), 1/sqrt(14)
            for m in (1, 2, 3, 6, 14):
                q = [x * math.sqrt(m) for x in r]
                if all(abs(x - round(x)) < 1e-12 for x in q) and sum(round(x) ** 2 for x in q) == m:
                    ints.append([int(round(x)) for x in q])
                    break
            else:
                raise ValueError(f"{arr} row {r} is not a unit Miller direction")
        out[arr] = ints
    return out


def main():
    lines = open(SRC).read().splitlines()
    blocks = {"fcc": (438, 515, 12), "bcc": (516, 593, 12), "roters": (604, 704, 12),
              "bcc12": (705, 811, 12), "bcc48": (812, 1207, 48)}
    data = {"source": "src/mod_crystals.f:438-515 (slip_type 1, fcc), :516-593 (2, bcc), "
                      ":604-704 (6, roters), :705-811 (7, bcc12), :812-1207 (8, bcc48)"}
    for name, (a, b, ns) in blocks.items():
        t = block(lines, a, b)
        assert len(t["bi"]) == ns, name
        data[name] = {"b": t["bi"], "n": t["ni"]}
    with open(os.path.join(HERE, "slip_tables.json"), "w") as f:
        json.dump(data, f, indent=None)
        f.write("\n")
    print(", ".join(f"{k} {len(v['b'])}" for k, v in data.items() if k != "source"),
          "systems written")


if __name__ == "__main__":
    main()
