#!/usr/bin/env python3
"""Write the water-monomer model tables the electrostatics and monomer terms need.

Reads the Fortran array initialisers of the reference's `module.f90` as text (nothing of
the reference is compiled or run) and writes them as plain number tables under
`dlpoly2-mbpol-heat_amd/data/`:

  dms_table.txt  84 rows "i1 i2 i3 coef": the Partridge-Schwenke dipole-moment-surface
                 expansion of module ps_type_dms (module.f90:765-812), used by qdetermine
                 (qdetermine.f90:153-181);
  ps_table.txt   245 rows "i1 i2 i3 c5z cbasis ccore crest": the Partridge-Schwenke
                 potential expansion of module ps_type_pes (module.f90:65-455), used by
                 ps_type_couple (ps_type_couple.f90:15-274);
  ps_mass.txt    9 rows "i1 i2 i3 cmass0": its mass-correction terms (module.f90:456-463).

The numbers are the model itself, so they must equal the reference's; only the file format
is this package's.  Usage:  python oracle/make_water_tables.py [/root/reference/SOURCE]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(HERE), "dlpoly2-mbpol-heat_amd", "data")


def module_text(src, name):
    m = re.search(r"^module\s+%s\b(.*?)^end module\s+%s" % (name, name), src, re.S | re.M | re.I)
    if not m:
        raise SystemExit(f"module {name} not found")
    return m.group(1)


def array(text, name, n, kind):
    m = re.search(r"::\s*%s\((?:1:)?%d\)\s*=\s*\(/(.*?)/\)" % (name, n), text, re.S | re.I)
    if not m:
        raise SystemExit(f"array {name}(1:{n}) not found")
    body = re.sub(r"!.*", "", m.group(1)).replace("&", " ")
    vals = [v for v in re.split(r"[\s,]+", body) if v]
    if len(vals) != n:
        raise SystemExit(f"{name}: expected {n} values, found {len(vals)}")
    if kind == "i":
        return [int(v) for v in vals]
    return [float(v.replace("D", "E").replace("d", "e")) for v in vals]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/SOURCE"
    src = open(os.path.join(ref, "module.f90")).read()
    os.makedirs(DATA, exist_ok=True)

    dms = module_text(src, "ps_type_dms")
    i1, i2, i3 = (array(dms, f"idx{k}", 84, "i") for k in (1, 2, 3))
    coef = array(dms, "coef", 84, "f")
    with open(os.path.join(DATA, "dms_table.txt"), "w") as f:
        for row in zip(i1, i2, i3, coef):
            f.write("%d %d %d %r\n" % row)

    pes = module_text(src, "ps_type_pes")
    j1, j2, j3 = (array(pes, f"idx{k}", 245, "i") for k in (1, 2, 3))
    cols = [array(pes, c, 245, "f") for c in ("c5z", "cbasis", "ccore", "crest")]
    with open(os.path.join(DATA, "ps_table.txt"), "w") as f:
        for row in zip(j1, j2, j3, *cols):
            f.write("%d %d %d %r %r %r %r\n" % row)
    m1, m2, m3 = (array(pes, f"idxm{k}", 9, "i") for k in (1, 2, 3))
    cm = array(pes, "cmass0", 9, "f")
    with open(os.path.join(DATA, "ps_mass.txt"), "w") as f:
        for row in zip(m1, m2, m3, cm):
            f.write("%d %d %d %r\n" % row)
    print("wrote", os.path.join(DATA, "dms_table.txt"), ", ps_table.txt and ps_mass.txt")


if __name__ == "__main__":
    main()
