"""GPU parity of the monomer bonds, PS coupling terms, TT dispersion and long-range
correction (libmbgpu.so, mbgpu_water_terms) against the oracle.

Bars (FP64): every energy and virial within 1e-10 of the oracle's (relative to the bond
energy, the largest term), stress, forces, per-atom energies and force-matrix
contractions within 1e-10 (max |diff| / max |value|).
"""
import dataclasses
import importlib

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def wp():
    return importlib.import_module("dlpoly2-mbpol-heat_amd.water_params")


def with_c8(p):
    return dataclasses.replace(p, c8=(5000.0, 1500.0, 300.0), a8=(3.1, 3.4, 3.7))


def run_gpu(pkg, box, p, ctx=None):
    nw = box.nw
    ctx = ctx or pkg.Context(nw, 4 * nw)
    xyz = o.site_xyz(box)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    f = np.zeros((3, 3 * nw))
    print(f"gpu water terms: nw={nw} rvdw={p.rvdw} ps={p.ps_monomer} disp={p.dispersion}", flush=True)
    r = ctx.water_terms(box.imcon, box.cell, p, x, y, z, f[0], f[1], f[2])
    r.update(ctx.water_details())
    r["f_host"] = f.T.copy()
    return ctx, r


def compare(g, r):
    scale = abs(r["eng_bnd"]) + abs(r["eng_ang"]) + abs(r["eng_srp"])
    for k in ("eng_bnd", "eng_ang", "eng_srp", "elrc", "vir_bnd", "vir_ang", "vir_srp", "virlrc"):
        assert abs(g[k] - r[k]) <= 1e-10 * scale, (k, g[k], r[k])
    assert rel(g["stress"], r["stress"]) <= 1e-10
    assert rel(g["f"], r["f"]) <= 1e-10
    assert np.array_equal(g["f"], g["f_host"])
    assert rel(g["atomic_energy"], r["atomic_energy"]) <= 1e-10
    assert rel(g["heat_w"], r["heat_w"]) <= 1e-10


def test_small_box_matches_oracle(pkg, waterbox, wp):
    box = waterbox.make_waterbox(16, 12.42, 4, seed=7)
    p = with_c8(wp.for_box(box.L))
    compare(run_gpu(pkg, box, p)[1], o.water_terms(box, p))


@pytest.mark.parametrize("terms", [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_each_term_alone(pkg, waterbox, wp, terms):
    box = waterbox.box_64()
    ps, disp, lr = terms
    p = with_c8(dataclasses.replace(wp.for_box(box.L), ps_monomer=ps, dispersion=disp, long_range=lr))
    if not ps:
        p = dataclasses.replace(p, bond_oh=(0, (0.0, 0.0, 0.0)), bond_hh=(0, (0.0, 0.0, 0.0)))
    g = run_gpu(pkg, box, p)[1]
    r = o.water_terms(box, p)
    scale = max(abs(r["eng_bnd"]) + abs(r["eng_ang"]) + abs(r["eng_srp"]), abs(r["elrc"]))
    for k in ("eng_bnd", "eng_ang", "eng_srp", "elrc", "vir_srp", "virlrc"):
        assert abs(g[k] - r[k]) <= 1e-10 * scale, k
    if ps or disp:
        assert rel(g["f"], r["f"]) <= 1e-10 and rel(g["atomic_energy"], r["atomic_energy"]) <= 1e-10


def test_box64_matches_oracle(pkg, waterbox, wp):
    box = waterbox.box_64()
    p = wp.for_box(box.L)
    compare(run_gpu(pkg, box, p)[1], o.water_terms(box, p))


def test_box512_matches_oracle_and_is_deterministic(pkg, waterbox, wp):
    box = waterbox.box_512()
    p = wp.for_box(box.L)
    ctx, a = run_gpu(pkg, box, p)
    compare(a, o.water_terms(box, p))
    _, b = run_gpu(pkg, box, p, ctx)
    for k in ("f", "atomic_energy", "heat_w", "stress"):
        assert np.array_equal(a[k], b[k]), k
    assert a["eng_srp"] == b["eng_srp"] and a["eng_ang"] == b["eng_ang"]


def test_heat_flux_includes_every_term(pkg, waterbox, wp):
    """J of MB-pol + electrostatics + water terms through the library vs the oracle."""
    ep = importlib.import_module("dlpoly2-mbpol-heat_amd.elec_params")
    box = waterbox.box_64()
    nw = box.nw
    pe = ep.for_box(box.L)
    pe.max_iter = 2000
    pw = wp.for_box(box.L)
    ctx = pkg.Context(nw, 4 * nw)
    f = np.zeros((3, box.natms))
    ctx.mbpol_forces(box.imcon, box.cell, 3 * nw, box.listttm2, box.x, box.y, box.z, f[0], f[1], f[2])
    ae_mb, hw_mb, _ = ctx.decomposition()
    xyz = o.site_xyz(box)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    d = np.zeros((3, 4 * nw))
    ctx.polar_electrostatics(box.imcon, box.cell, pe, x.copy(), y.copy(), z.copy(), d[0], d[1], d[2])
    ctx.water_terms(box.imcon, box.cell, pw, x, y, z)
    vel = box.vel
    ek = 0.5 * box.mass * (vel ** 2).sum(1)
    jk, jp, jt = ctx.heat_flux(vel[:, 0], vel[:, 1], vel[:, 2], ek)
    re = o.elec_forces(box, pe)
    rw = o.water_terms(box, pw)
    nr = 3 * nw
    hw = hw_mb.reshape(box.natms, 9).copy()
    hw[:nr] += re["heat_w"].reshape(nr, 9) + rw["heat_w"].reshape(nr, 9)
    e = ae_mb.copy()
    e[:nr] += re["atomic_energy"] + rw["atomic_energy"]
    s = np.zeros((box.natms, 9))
    s[:nr] = re["s_ew1p"]
    rk, rp, rt = o.heat_flux_tensor(box, hw.reshape(-1), vel, ek, e, s.reshape(-1).copy())
    assert rel(jk, rk) <= 1e-10 and rel(jp, rp) <= 1e-10 and rel(jt, rt) <= 1e-10
    # the water terms alone change J: they are really included
    ctx2 = pkg.Context(nw, 4 * nw)
    ctx2.mbpol_forces(box.imcon, box.cell, 3 * nw, box.listttm2, box.x, box.y, box.z, f[0], f[1], f[2])
    ctx2.polar_electrostatics(box.imcon, box.cell, pe, x.copy(), y.copy(), z.copy(), d[0], d[1], d[2])
    jk2, jp2, _ = ctx2.heat_flux(vel[:, 0], vel[:, 1], vel[:, 2], ek)
    assert rel(jp2, jp) > 1e-6


def test_parameter_errors(pkg, waterbox, wp):
    box = waterbox.make_waterbox(8, 12.42, 2, seed=3)
    ctx = pkg.Context(box.nw, 4 * box.nw)
    a = [np.zeros(3 * box.nw) for _ in range(3)]
    with pytest.raises(pkg.MbgpuError, match="bond_key"):
        ctx.water_terms(1, box.cell, dataclasses.replace(wp.for_box(box.L), bond_hh=(5, (1.0, 1.0, 0.0))), *a)
    with pytest.raises(pkg.MbgpuError, match="rvdw"):
        ctx.water_terms(1, box.cell, dataclasses.replace(wp.for_box(box.L), rcut=5.0), *a)
    with pytest.raises(pkg.MbgpuError, match="imcon"):
        ctx.water_terms(4, box.cell, wp.for_box(box.L), *a)
