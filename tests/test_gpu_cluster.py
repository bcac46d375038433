"""Isolated clusters (imcon 0) on the GPU against the oracle.

The reference has no isolated cluster: image_vectors stops for imcon 0 (mbpol.f90:1357) and
its polarizable electrostatics need the Ewald sum (forces.f:271-273).  The semantics here
are this library's: no images; the electrostatics are direct sums over every site pair (no
Ewald split: alpha = 0, so the real-space kernels are the bare ones, the -erf parts vanish
as ewald3p.f sets them, and there is no reciprocal sum or self term); no long-range
dispersion tail (lrcorrect.f:114).  The oracle implements the same; its direct sum is
pinned on the CPU as the open-boundary limit of the Ewald sum
(test_oracle_elec.py::test_cluster_direct_sum_is_the_open_boundary_limit_of_ewald).

* A 40-water cluster: MB-pol (lists bit-exact, energy, virial, forces), the electrostatics
  with both solvers and the water terms within 1e-10 of the oracle, per-atom arrays too.
* The whole evaluation (engine.Slot) with the heat flux against the oracle's sum.
* Arguments that make no sense for a cluster are errors.
"""
import importlib

import numpy as np
import pytest

import oracle_lib as o
from test_gpu_elec import run_gpu

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def cluster(waterbox):
    return waterbox.make_cluster(40, seed=3)


def test_mbpol_on_a_cluster(pkg, cluster):
    box = cluster
    ctx = pkg.Context(box.nw, box.natms)
    f = np.zeros((3, box.natms))
    st = np.zeros(9)
    print("cluster of 40 waters: MB-pol", flush=True)
    e, v = ctx.mbpol_forces(0, box.cell, 3 * box.nw, box.listttm2, box.x, box.y, box.z, f[0], f[1], f[2],
                            0.0, 0.0, st)
    r = o.mbpol_forces(box)
    d, t = ctx.lists()
    assert sorted(map(tuple, d.tolist())) == sorted(map(tuple, r["dimers"].tolist()))
    assert sorted(map(tuple, t.tolist())) == sorted(map(tuple, r["trimers"].tolist()))
    assert abs(e - r["eng"]) <= 1e-10 * abs(r["eng"]) and abs(v - r["vir"]) <= 1e-10 * abs(r["vir"])
    assert rel(st, r["stress"]) <= 1e-10 and rel(f.T, r["f"]) <= 1e-10


@pytest.mark.parametrize("solver", [0, 1])
def test_electrostatics_on_a_cluster(pkg, cluster, solver):
    p = pkg.elec_params.for_cluster(cluster.L)
    p.max_iter = 2000
    p.solver = solver
    r = o.elec_forces(cluster, p)
    assert r["eng_ew1p"] == 0.0 and r["eng_ew3p"] == 0.0 and r["vir_ew1p"] == 0.0
    _, g = run_gpu(pkg, cluster, p)
    # no reciprocal part: energies on the scale of the real-space sum
    scale = abs(r["eng_ew2p"]) + abs(r["eng_polar"])
    for k in ("eng_ew1p", "eng_ew2p", "eng_ew3p", "eng_polar", "eng"):
        assert abs(g[k] - r[k]) <= 1e-10 * scale, k
    assert abs(g["vir"] - r["vir"]) <= 1e-10 * abs(r["vir"])
    assert rel(g["xyz"], r["xyz"]) == 0.0
    assert rel(g["chge"], r["chge"]) <= 1e-12
    for k in ("stress", "dip", "vesp", "f", "atomic_energy", "heat_w"):
        assert rel(g[k], r[k]) <= 1e-10, k
    assert not np.any(g["s_ew1p"]) and not np.any(r["s_ew1p"])
    if solver == 0:   # the oracle iterates as the SOR solver does; CG takes fewer steps
        assert abs(g["niter"] - r["niter"]) <= 1
    else:
        assert 0 < g["niter"] < r["niter"]


def test_water_terms_on_a_cluster(pkg, cluster):
    box = cluster
    p = pkg.water_params.for_cluster(box.L)
    ctx = pkg.Context(box.nw, box.natms)
    xyz = o.site_xyz(box)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    f = np.zeros((3, 3 * box.nw))
    w = ctx.water_terms(0, box.cell, p, x, y, z, f[0], f[1], f[2])
    rw = o.water_terms(box, p)
    assert w["elrc"] == 0.0 and w["virlrc"] == 0.0 and rw["elrc"] == 0.0
    scale = abs(rw["eng_bnd"]) + abs(rw["eng_ang"]) + abs(rw["eng_srp"])
    for k in ("eng_bnd", "eng_ang", "eng_srp", "vir_bnd", "vir_ang", "vir_srp"):
        assert abs(w[k] - rw[k]) <= 1e-10 * scale, k
    assert rel(w["stress"], rw["stress"]) <= 1e-10 and rel(f.T, rw["f"]) <= 1e-10
    wd = ctx.water_details()
    for k in ("atomic_energy", "heat_w"):
        assert rel(wd[k], rw[k]) <= 1e-10, k


def test_whole_evaluation_of_a_cluster(pkg, cluster):
    engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
    box = cluster
    pe, pw = pkg.elec_params.for_cluster(box.L), pkg.water_params.for_cluster(box.L)
    pe.max_iter = 2000
    slot = engine.Slot(box, solver=0, elec=pe, water=pw)
    r = slot.evaluate()
    rm, re, rw = o.mbpol_forces(box), o.elec_forces(box, pe), o.water_terms(box, pw)
    e_ref = rm["eng"] + re["eng"] + rw["eng_bnd"] + rw["eng_ang"] + rw["eng_srp"]
    assert abs(r["eng"] - e_ref) <= 1e-10 * abs(rw["eng_bnd"])
    nr = box.nreal
    f_ref = rm["f"][:nr] + re["f"] + rw["f"]
    assert rel(slot.forces(), f_ref) <= 1e-10
    vel = box.vel
    ek = 0.5 * box.mass * (vel ** 2).sum(1)
    hw = rm["heat_w"].reshape(box.natms, 9).copy()
    hw[:nr] += re["heat_w"].reshape(nr, 9) + rw["heat_w"].reshape(nr, 9)
    e = rm["atomic_energy"].copy()
    e[:nr] += re["atomic_energy"] + rw["atomic_energy"]
    s = np.zeros((box.natms, 9))
    s[:nr] = re["s_ew1p"]
    _, _, jt = o.heat_flux_tensor(box, hw.reshape(-1), vel, ek, e, s.reshape(-1).copy())
    assert rel(r["jtot"], jt) <= 1e-10


def test_cluster_argument_errors(pkg, cluster):
    box = cluster
    ctx = pkg.Context(box.nw, box.natms)
    xyz = o.site_xyz(box)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    d = np.zeros((3, box.natms))
    f = np.zeros((3, box.natms))
    p = pkg.elec_params.for_cluster(box.L)
    p.alpha = 0.3
    with pytest.raises(pkg.MbgpuError, match="alpha = 0"):
        ctx.polar_electrostatics(0, box.cell, p, x, y, z, d[0], d[1], d[2], f[0], f[1], f[2])
    ctx.set_elec_reciprocal("spme", 6, (32, 32, 32))
    with pytest.raises(pkg.MbgpuError, match="periodic"):
        ctx.polar_electrostatics(0, box.cell, pkg.elec_params.for_cluster(box.L), x, y, z, d[0], d[1], d[2],
                                 f[0], f[1], f[2])
