#!/usr/bin/env python3
"""Derive the MB-pol polynomial data tables from the reference (run once, here only).

Test/data infrastructure: needs the reference tree and the compiled reference leaves
(``make -C oracle ref``).  It writes the tables the product's kernel generator reads:

  <pkg>/data/poly2b_monomials.txt  one line per monomial "basis v1 [v2 [v3 [v4]]]"
  <pkg>/data/poly3b_monomials.txt
  <pkg>/data/poly2b_coeffs.txt     the fitted coefficient vector a[1153]
  <pkg>/data/poly3b_coeffs.txt     the fitted coefficient vector a[1163]

Monomials come from black-box probing of poly_2b_v6x::eval / poly_3b_v2x::eval
(oracle/extract_pip_tables.c, which also verifies the table against eval()).
Coefficients are the numeric table ``the_poly[]`` of mbpol-2b-poly.cpp:688-1842 and
mbpol-3b-poly.cpp:318-1482, read as data.
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
PKG = os.path.join(REPO, "dlpoly2-mbpol-heat_amd")
REF = os.environ.get("MBPOL_REF", "/root/reference/SOURCE")


def coeffs_from_reference(path, n):
    text = open(path).read()
    m = re.search(r"the_poly\[\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise SystemExit(f"no coefficient table in {path}")
    body = re.sub(r"//[^\n]*", "", m.group(1))
    vals = [float(t) for t in body.replace("\n", " ").split(",") if t.strip()]
    if len(vals) != n:
        raise SystemExit(f"{path}: expected {n} coefficients, found {len(vals)}")
    return vals


def main():
    lib = os.path.join(HERE, "_ref", "libmbpol_ref.so")
    if not os.path.exists(lib):
        raise SystemExit("build the reference leaves first: make -C oracle ref")
    exe = os.path.join(HERE, "_ref", "extract_pip_tables")
    subprocess.check_call(["gcc", "-O2", "-fopenmp", "-o", exe,
                           os.path.join(HERE, "extract_pip_tables.c"),
                           "-L" + os.path.dirname(lib), "-lmbpol_ref",
                           "-Wl,-rpath," + os.path.dirname(lib), "-lm"])
    os.makedirs(os.path.join(PKG, "data"), exist_ok=True)
    for which, n, src in ((2, 1153, "mbpol-2b-poly.cpp"), (3, 1163, "mbpol-3b-poly.cpp")):
        out = subprocess.check_output([exe, str(which)])
        rows = []
        for line in out.decode().splitlines():
            t = line.split()
            if t[1] != "1":
                raise SystemExit("unexpected non-unit monomial coefficient")
            rows.append(" ".join([t[0]] + t[2:]))
        with open(os.path.join(PKG, "data", f"poly{which}b_monomials.txt"), "w") as f:
            f.write("\n".join(rows) + "\n")
        a = coeffs_from_reference(os.path.join(REF, src), n)
        with open(os.path.join(PKG, "data", f"poly{which}b_coeffs.txt"), "w") as f:
            f.write("\n".join(repr(v) for v in a) + "\n")
        print(f"poly{which}b: {len(rows)} monomials, {n} coefficients", file=sys.stderr)


if __name__ == "__main__":
    main()
