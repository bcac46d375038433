"""Forces against finite differences of the energies, per term, on the oracle (CPU).

Central differences (h = 1e-5 A) of the oracle's energies against its forces at a few
coordinates of a 16-water box:
* MB-pol (2B + 3B) and the water terms (PS monomer, TT dispersion): within 1e-8 of the
  largest force -- the forces are the exact gradients.
* Electrostatics: within 1e-5 of the largest force, and no closer than 1e-7: the
  reference's erfc is the Numerical Recipes approximation with single-precision literals
  (erfcc.f, fractional error < 1.2e-7), and its Ewald real-space and intramolecular terms
  take the force from the recursion B_1 = (B_0 + 2 alpha / sqrt(pi) exp(-a^2)) / r^2, which
  is the derivative of the exact erfc, not of the approximation (ewald3p.f:104-114,
  erfcgenp.f).  The mismatch (about 2e-6 of the largest force) is a property of the
  reference's arithmetic, which this repository keeps for parity; with the exact erfc in
  its place (the oracle's second build) it falls within 1e-8, the differences' own noise.
  A dropped term, a wrong sign or a wrong factor would give a mismatch of order one.
"""
import dataclasses
import importlib

import numpy as np

import oracle_lib as o

H = 1e-5
COORDS = ((1, 0), (13, 1), (28, 2))


def moved(wb, box, k, a, h):
    b = wb.WaterBox(box.nw, box.L, box.xyz.copy(), box.vel.copy(), box.mass.copy(), box.listttm2.copy())
    b.xyz[k, a] += h
    return b


def test_forces_are_gradients_of_the_energies():
    wb = importlib.import_module("dlpoly2-mbpol-heat_amd.waterbox")
    ep = importlib.import_module("dlpoly2-mbpol-heat_amd.elec_params")
    wp = importlib.import_module("dlpoly2-mbpol-heat_amd.water_params")
    box = wb.make_waterbox(16, 12.42, 4, seed=7)
    pe = ep.ElecParams(rcut=6.2, alpha=0.72, kmax=(12, 12, 12), max_iter=2000)
    pw = dataclasses.replace(wp.for_box(box.L), c8=(5000.0, 1500.0, 300.0), a8=(3.1, 3.4, 3.7))

    def terms(b):
        m, e, w = o.mbpol_forces(b), o.elec_forces(b, pe), o.water_terms(b, pw)
        x = o.elec_forces(b, pe, exact_erfc=True)
        return {"mbpol": (m["eng"], m["f"][:b.nreal]), "elec": (e["eng"], e["f"]),
                "water": (w["eng_bnd"] + w["eng_ang"] + w["eng_srp"], w["f"]), "elec_exact": (x["eng"], x["f"])}

    base = terms(box)
    worst = {k: 0.0 for k in base}
    for k, a in COORDS:
        p, m = terms(moved(wb, box, k, a, H)), terms(moved(wb, box, k, a, -H))
        for t in base:
            fd = -(p[t][0] - m[t][0]) / (2 * H)
            worst[t] = max(worst[t], abs(base[t][1][k, a] - fd))
    rel = {t: worst[t] / np.abs(base[t][1]).max() for t in base}
    print("max |F - FD| / max |F|:", {t: f"{v:.2e}" for t, v in rel.items()})
    assert rel["mbpol"] <= 1e-8 and rel["water"] <= 1e-8
    assert 1e-7 <= rel["elec"] <= 1e-5
    assert rel["elec_exact"] <= 1e-8
