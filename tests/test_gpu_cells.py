"""Non-cubic cells (imcon 2, rectangular; imcon 3, parallelepiped) through the whole force
path on the GPU against the oracle with the same cell: MB-pol (lists bit-exact, energy,
virial, stress, forces), polarizable electrostatics and the water terms (energies, virial,
stress, forces) within 1e-10.  The 64-water box is mapped affinely onto each cell (a few
per cent of shear and stretch); the cutoffs fit the cells' perpendicular widths.

The 512-water box mapped the same way with 5 A cutoffs holds four link cells per axis; with
the cells taken from 1 024 particles on (MBGPU_CELLS_MIN; by default they are used on
larger boxes) its electrostatics neighbour rows and dispersion rows come from the 27 cells
around each particle: the same 1e-10 bars, against the oracle's all-pairs scans."""
import dataclasses

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cell_box(waterbox, imcon, base=None):
    base = base or waterbox.box_64()
    L = base.L
    if imcon == 2:
        H = np.diag([1.06 * L, 0.96 * L, 1.0 * L])
    else:
        H = np.array([[L, 0.0, 0.0], [0.10 * L, 0.97 * L, 0.0], [0.05 * L, -0.07 * L, 1.02 * L]])
    s = base.xyz / L
    xyz = s @ H
    box = waterbox.WaterBox(base.nw, L, xyz, base.vel.copy(), base.mass.copy(), base.listttm2.copy())
    box.cell = np.ascontiguousarray(H.reshape(-1))   # a, b, c as DL_POLY's cell(1:9)
    box.imcon = imcon
    return box


@pytest.mark.parametrize("imcon", [2, 3])
def test_whole_path_on_non_cubic_cells(pkg, waterbox, imcon):
    box = cell_box(waterbox, imcon)
    nw, n = box.nw, box.natms
    print(f"imcon {imcon}: 64 waters, cell {np.round(box.cell, 3).tolist()}", flush=True)
    # MB-pol
    ctx = pkg.Context(nw, n)
    f = np.zeros((3, n))
    st = np.zeros(9)
    e, v = ctx.mbpol_forces(box.imcon, box.cell, 3 * nw, box.listttm2, box.x, box.y, box.z, f[0], f[1], f[2],
                            0.0, 0.0, st)
    r = o.mbpol_forces(box)
    d, t = ctx.lists()
    assert sorted(map(tuple, d.tolist())) == sorted(map(tuple, r["dimers"].tolist()))
    assert sorted(map(tuple, t.tolist())) == sorted(map(tuple, r["trimers"].tolist()))
    assert abs(e - r["eng"]) <= 1e-10 * abs(r["eng"]) and abs(v - r["vir"]) <= 1e-10 * abs(r["vir"])
    assert rel(st, r["stress"]) <= 1e-10 and rel(f.T, r["f"]) <= 1e-10
    # electrostatics
    pe = pkg.elec_params.ElecParams(rcut=5.0, alpha=0.65, kmax=(8, 8, 8), max_iter=2000)
    xyz = o.site_xyz(box)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    dd = np.zeros((3, n))
    fe = np.zeros((3, n))
    g = ctx.polar_electrostatics(box.imcon, box.cell, pe, x, y, z, dd[0], dd[1], dd[2], fe[0], fe[1], fe[2])
    re = o.elec_forces(box, pe)
    assert abs(g["eng"] - re["eng"]) <= 1e-10 * abs(re["eng_ew1p"])
    assert abs(g["vir"] - re["vir"]) <= 1e-10 * abs(re["vir_ew1p"])
    assert rel(g["stress"], re["stress"]) <= 1e-10
    assert rel(fe.T[:3 * nw], re["f"]) <= 1e-10
    # water terms
    pw = dataclasses.replace(pkg.water_params.for_box(box.L), rvdw=5.0, rcut=5.0)
    fw = np.zeros((3, 3 * nw))
    w = ctx.water_terms(box.imcon, box.cell, pw, x, y, z, fw[0], fw[1], fw[2])
    rw = o.water_terms(box, pw)
    scale = abs(rw["eng_bnd"]) + abs(rw["eng_ang"]) + abs(rw["eng_srp"])
    for k in ("eng_bnd", "eng_ang", "eng_srp", "elrc", "vir_bnd", "vir_ang", "vir_srp", "virlrc"):
        assert abs(w[k] - rw[k]) <= 1e-10 * scale, k
    assert rel(w["stress"], rw["stress"]) <= 1e-10 and rel(fw.T, rw["f"]) <= 1e-10


@pytest.mark.parametrize("imcon", [2, 3])
def test_link_cells_on_non_cubic_cells(pkg, waterbox, imcon, monkeypatch):
    # the searches take the cells from 1 024 particles on (by default only on larger boxes,
    # where they beat the scans)
    monkeypatch.setenv("MBGPU_CELLS_MIN", "1024")
    box = cell_box(waterbox, imcon, waterbox.box_512())
    nw, n = box.nw, box.natms
    print(f"imcon {imcon}: 512 waters through link cells, cell {np.round(box.cell, 3).tolist()}", flush=True)
    ctx = pkg.Context(nw, n)
    pe = pkg.elec_params.ElecParams(rcut=5.0, alpha=0.65, kmax=(8, 8, 8), max_iter=2000)
    xyz = o.site_xyz(box)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    dd = np.zeros((3, n))
    fe = np.zeros((3, n))
    g = ctx.polar_electrostatics(box.imcon, box.cell, pe, x, y, z, dd[0], dd[1], dd[2], fe[0], fe[1], fe[2])
    re = o.elec_forces(box, pe)
    assert abs(g["eng"] - re["eng"]) <= 1e-10 * abs(re["eng_ew1p"])
    assert abs(g["vir"] - re["vir"]) <= 1e-10 * abs(re["vir_ew1p"])
    assert rel(g["stress"], re["stress"]) <= 1e-10
    assert rel(fe.T[:3 * nw], re["f"]) <= 1e-10
    det = ctx.elec_details()
    for k in ("atomic_energy", "heat_w", "s_ew1p", "vesp"):
        assert rel(det[k], re[k]) <= 1e-10, k
    pw = dataclasses.replace(pkg.water_params.for_box(box.L), rvdw=5.0, rcut=5.0)
    fw = np.zeros((3, 3 * nw))
    w = ctx.water_terms(box.imcon, box.cell, pw, x, y, z, fw[0], fw[1], fw[2])
    rw = o.water_terms(box, pw)
    scale = abs(rw["eng_bnd"]) + abs(rw["eng_ang"]) + abs(rw["eng_srp"])
    for k in ("eng_bnd", "eng_ang", "eng_srp", "elrc", "vir_bnd", "vir_ang", "vir_srp", "virlrc"):
        assert abs(w[k] - rw[k]) <= 1e-10 * scale, k
    assert rel(w["stress"], rw["stress"]) <= 1e-10 and rel(fw.T, rw["f"]) <= 1e-10
    wd = ctx.water_details()
    for k in ("atomic_energy", "heat_w"):
        assert rel(wd[k], rw[k]) <= 1e-10, k
