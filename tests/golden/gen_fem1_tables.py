#!/usr/bin/env python3
"""Extract the 12x12 element-matrix tables of the first-order stabilized hex
solvers (src/fem, src/fem_abm) from the reference sources into data files.

The tables are literals in the reference, written as quotients of decimals
(e.g. "4.0 / 3.0").  Each entry is turned into an exact fraction; the double
the reference uses is the correctly rounded value of that fraction, which is
also what num/den gives in IEEE double arithmetic.

  src/fem/fem_simulation.cpp:133-260      use_ofem = 0: A, _A, B, _B, C, _C
  src/fem/fem_simulation.cpp:266-406      use_ofem = 1: A, _A, B, _B, C, _C
  src/fem_abm/fem_simulation.cpp:146-189  A, _A (the active definitions)

Outputs (both committed):
  tests/golden/fem1_tables.json         numerator/denominator pairs (fixture)
  emsolvers_amd/csrc/fem1_tables.inc    the same as C integer tables

Usage: python tests/golden/gen_fem1_tables.py [/root/reference]
"""
import json
import math
import os
import re
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "emsolvers_amd", "csrc")

NAMES = ["Amat", "_Amat", "Bmat", "_Bmat", "Cmat", "_Cmat"]


def parse_tables(path):
    """Every active `std::array<std::array<double, 12>, 12> NAME = {...}`
    literal in file order, as (name, 12x12 Fractions)."""
    lines = open(path, encoding="utf-8").read().splitlines()
    out = []
    decl = re.compile(r"^\s*std::array<std::array<double, 12>, 12> (\w+) = \{")
    i = 0
    while i < len(lines):
        m = decl.match(lines[i])
        if not m:
            i += 1
            continue
        name = m.group(1)
        body = []
        depth = lines[i].count("{") - lines[i].count("}")
        i += 1
        while depth > 0 and i < len(lines):
            s = lines[i].split("//", 1)[0]
            body.append(s)
            depth += s.count("{") - s.count("}")
            i += 1
        text = " ".join(body).replace("{", " ").replace("}", " ").replace(";", " ")
        vals = [v.strip() for v in text.split(",") if v.strip()]
        ent = []
        for v in vals:
            if "/" in v:
                a, b = v.split("/")
                ent.append(Fraction(a.strip()) / Fraction(b.strip()))
            else:
                ent.append(Fraction(v))
        if len(ent) != 144:
            raise SystemExit("%s: table %s has %d entries" % (path, name, len(ent)))
        out.append((name, [ent[r * 12:(r + 1) * 12] for r in range(12)]))
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    fem = parse_tables(os.path.join(ref, "src", "fem", "fem_simulation.cpp"))
    abm = parse_tables(os.path.join(ref, "src", "fem_abm", "fem_simulation.cpp"))
    if [n for n, _ in fem] != NAMES * 2:
        raise SystemExit("unexpected table sequence in src/fem: %s" % [n for n, _ in fem])
    if [n for n, _ in abm] != NAMES[:2]:
        raise SystemExit("unexpected table sequence in src/fem_abm: %s" % [n for n, _ in abm])
    sets = {
        "fem_ofem0": dict(fem[:6]),
        "fem_ofem1": dict(fem[6:]),
        "fem_abm": dict(abm),
    }
    js = {k: {n: [["%d/%d" % (x.numerator, x.denominator) for x in row] for row in t] for n, t in v.items()}
          for k, v in sets.items()}
    with open(os.path.join(ROOT, "tests", "golden", "fem1_tables.json"), "w") as f:
        json.dump(js, f, indent=0)
        f.write("\n")
    # C include: one int numerator table per matrix and a common denominator
    lines = ["// Element-matrix tables of the first-order stabilized hex solvers,",
             "// generated by tests/golden/gen_fem1_tables.py from",
             "// src/fem/fem_simulation.cpp:133-406 and src/fem_abm/fem_simulation.cpp:146-189.",
             "// value = num[r*12+c] / den (IEEE double division).", ""]
    for key, tabs in sets.items():
        for n in NAMES:
            if n not in tabs:
                continue
            t = tabs[n]
            den = 1
            for row in t:
                for x in row:
                    den = den * x.denominator // math.gcd(den, x.denominator)
            nums = [int(x * den) for row in t for x in row]
            cname = "k%s_%s" % (key, n.replace("_", "s") if n.startswith("_") else n)
            lines.append("static const int %s_den = %d;" % (cname, den))
            lines.append("static const signed short %s[144] = {" % cname)
            for r in range(12):
                lines.append("    " + ", ".join("%d" % v for v in nums[r * 12:(r + 1) * 12]) + ",")
            lines.append("};")
    with open(os.path.join(CSRC, "fem1_tables.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote tests/golden/fem1_tables.json and emsolvers_amd/csrc/fem1_tables.inc")


if __name__ == "__main__":
    main()
