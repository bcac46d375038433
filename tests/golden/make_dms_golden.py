#!/usr/bin/env python3
"""Write golden vectors of the DMS charges and their forces from the reference's own code.

Run once, in a container that holds the reference tree: ``make -C oracle ref`` compiles
qdetermine.f90 and qdforce.f90 with the modules and helpers they need (module.f90,
multibead.f90, error.f, invert.f, images.f, gdsum.f, heatcurrent.f90) where they lie, with
amdflang in the reference's serial mode, plus the harness oracle/ref_dms.f90,
into oracle/_ref/libdms_ref.so.  Output: tests/golden/dms_reference.npz, inputs and the
reference's outputs only:

  xyz    (nset, 4*nw, 3)  O, H1, H2 of each water, then the M sites (Angstrom)
  cell   (nset, 9)        the cell of each set (Fortran cell(1..9))
  imcon  (nset,)          1, 2 or 3
  gamma  (nset,)          the M-site fraction
  chge   (nset, 4*nw)     charges qdetermine assigns (O stays 0, H1, H2, M)
  dp     (nset, nw, 6)    dp1/dr1, dp1/dr2, dp1/dcos, dp2/dr1, dp2/dr2, dp2/dcos
  moldip (nset, nw, 3)    the reference's molecular dipole of the assigned charges
  chkdip (nset, nw, 3)    its PS dipole p1 r1 + p2 r2 (qdetermine.f90:262-305)

and the reference's qdforce (qdforce.f90, built with HEAT_CURRENT and without HEAT_STRESS as
the reference's Makefile sets it) for 9 sets of 8 waters with random site potentials:

  qd_xyz (nset, 32, 3), qd_cell (nset, 9), qd_imcon, qd_gamma, qd_vesp (nset, 32)
  qd_f      (nset, 32, 3)    charge-derivative forces at the sites (M not redistributed)
  qd_stress (nset, 9), qd_vir (nset,)
  qd_sqa    (nset, 9)        sum over atoms of atomic_stress_qd
  qd_fm     (nset, 3, 32, 32) force_matrix(i, j, k) stored [k][j][i]

Geometries: waters with O-H lengths in [0.75, 1.35] A and HOH angles in [70, 150] deg,
random orientations, placed so that some hydrogens sit across the cell boundary (their
bonds are found through the minimum image), in a cubic, an orthorhombic and a
parallelepiped cell, with gamma = 0.426706882 (MB-pol) and two other values.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libdms_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(_dp)


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_set(rng, nw, cell, gamma):
    cv = np.asarray(cell).reshape(3, 3)           # rows: cell vectors
    xyz = np.zeros((4 * nw, 3))
    for i in range(nw):
        r1, r2 = rng.uniform(0.75, 1.35, 2)
        th = np.radians(rng.uniform(70.0, 150.0))
        R = rotation(rng)
        # O anywhere in the cell, near a face for every third water
        s = rng.uniform(-0.5, 0.5, 3)
        if i % 3 == 0:
            s[i % 3] = 0.5 - 1e-3
        O = s @ cv
        h1 = O + R @ np.array([r1, 0.0, 0.0])
        h2 = O + R @ np.array([r2 * np.cos(th), r2 * np.sin(th), 0.0])
        # wrap the hydrogens back into the cell, so that their bonds cross the boundary
        for h in (h1, h2):
            f = h @ np.linalg.inv(cv)
            h -= np.round(f) @ cv
        xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2] = O, h1, h2
        d1 = h1 - O
        d2 = h2 - O
        d1 -= np.round(d1 @ np.linalg.inv(cv)) @ cv
        d2 -= np.round(d2 @ np.linalg.inv(cv)) @ cv
        xyz[3 * nw + i] = O + 0.5 * gamma * (d1 + d2)
    return xyz


def main():
    if not os.path.exists(LIB):
        raise SystemExit(f"{LIB} is missing: make -C oracle ref")
    lib = ctypes.CDLL(LIB)
    lib.ref_dms.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double] + [_dp] * 5
    rng = np.random.Generator(np.random.PCG64(20261203))
    nw = 24
    cells = [(1, [9.5, 0, 0, 0, 9.5, 0, 0, 0, 9.5]),
             (2, [8.0, 0, 0, 0, 11.0, 0, 0, 0, 9.0]),
             (3, [9.0, 0.0, 0.0, 1.5, 9.5, 0.0, -1.0, 0.8, 10.0])]
    out = {k: [] for k in ("xyz", "cell", "imcon", "gamma", "chge", "dp", "moldip", "chkdip")}
    for gamma in (0.426706882, 0.2, 0.65):
        for imcon, cell in cells:
            xyz = np.ascontiguousarray(water_set(rng, nw, cell, gamma))
            c = np.asarray(cell, float)
            chge, dp = np.zeros(4 * nw), np.zeros((nw, 6))
            md, cd = np.zeros((nw, 3)), np.zeros((nw, 3))
            lib.ref_dms(nw, imcon, P(c), gamma, P(xyz), P(chge), P(dp), P(md), P(cd))
            for k, v in (("xyz", xyz), ("cell", c), ("imcon", imcon), ("gamma", gamma), ("chge", chge),
                         ("dp", dp), ("moldip", md), ("chkdip", cd)):
                out[k].append(v)
    lib.ref_qd.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double] + [_dp] * 7
    nq = 8
    qd = {k: [] for k in ("qd_xyz", "qd_cell", "qd_imcon", "qd_gamma", "qd_vesp", "qd_f", "qd_stress", "qd_vir",
                          "qd_sqa", "qd_fm")}
    for gamma in (0.426706882, 0.2, 0.65):
        for imcon, cell in cells:
            xyz = np.ascontiguousarray(water_set(rng, nq, cell, gamma))
            c = np.asarray(cell, float)
            n = 4 * nq
            vesp = rng.normal(size=n) * 50.0
            f, st, sqa, fm = np.zeros((n, 3)), np.zeros(9), np.zeros(9), np.zeros((3, n, n))
            vir = ctypes.c_double()
            lib.ref_qd(nq, imcon, P(c), gamma, P(xyz), P(vesp), P(f), P(st), ctypes.byref(vir), P(sqa), P(fm))
            for k, v in (("qd_xyz", xyz), ("qd_cell", c), ("qd_imcon", imcon), ("qd_gamma", gamma),
                         ("qd_vesp", vesp), ("qd_f", f), ("qd_stress", st), ("qd_vir", vir.value), ("qd_sqa", sqa),
                         ("qd_fm", fm)):
                qd[k].append(v)
    out.update(qd)
    np.savez_compressed(os.path.join(HERE, "dms_reference.npz"), **{k: np.array(v) for k, v in out.items()})
    print("wrote", len(out["xyz"]), "sets of", nw, "waters and", len(qd["qd_f"]), "qdforce sets of", nq)


if __name__ == "__main__":
    main()
