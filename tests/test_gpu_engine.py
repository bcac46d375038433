"""Complete force evaluations through engine.Slot, alone and concurrently.

* One slot on the 64-water box: total energy, virial, forces and heat flux equal the sum of
  the oracle's MB-pol, electrostatics and water terms within 1e-10.
* Several slots driven concurrently from host threads (their kernels overlap on the GPU):
  every result is bit-identical to the same slot evaluated alone.
* The electrostatics on the context's second stream from a helper thread, beside MB-pol and
  the water terms (Slot overlap), give bit-identical results to the three parts in order.
* The FP32-polynomial mode against the mixed-precision bar (expected to fail: see
  FP32_BAR_REASON), and its FP64 parts unchanged.
* Electrostatics neighbour rows that start too small (MBGPU_ELEC_ROWS): the evaluation
  grows them and repeats from the same start dipoles, bit-identical to a normal context.
"""
import importlib

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(pkg):
    return importlib.import_module("dlpoly2-mbpol-heat_amd.engine")


def test_slot_matches_oracle_sum(pkg, waterbox, engine):
    box = waterbox.box_64()
    pe = pkg.elec_params.for_box(box.L)
    pe.max_iter = 2000
    slot = engine.Slot(box, solver=0, elec=pe)
    print("slot evaluate, 64 waters", flush=True)
    r = slot.evaluate()
    f = slot.forces()
    rm = o.mbpol_forces(box)
    re = o.elec_forces(box, pe)
    rw = o.water_terms(box, pkg.water_params.for_box(box.L))
    e_ref = rm["eng"] + re["eng"] + rw["eng_bnd"] + rw["eng_ang"] + rw["eng_srp"] + rw["elrc"]
    v_ref = rm["vir"] + re["vir"] + rw["vir_bnd"] + rw["vir_ang"] + rw["vir_srp"] + rw["virlrc"]
    scale = abs(rw["eng_bnd"])
    assert abs(r["eng"] - e_ref) <= 1e-10 * scale
    assert abs(r["vir"] - v_ref) <= 1e-10 * abs(re["vir_ew1p"])
    nr = box.nreal
    f_ref = rm["f"][:nr] + re["f"] + rw["f"]
    assert np.abs(f - f_ref).max() <= 1e-10 * np.abs(f_ref).max()
    # heat flux over every term
    vel = box.vel
    ek = 0.5 * box.mass * (vel ** 2).sum(1)
    hw = rm["heat_w"].reshape(box.natms, 9).copy()
    hw[:nr] += re["heat_w"].reshape(nr, 9) + rw["heat_w"].reshape(nr, 9)
    e = rm["atomic_energy"].copy()
    e[:nr] += re["atomic_energy"] + rw["atomic_energy"]
    s = np.zeros((box.natms, 9))
    s[:nr] = re["s_ew1p"]
    _, _, jt = o.heat_flux_tensor(box, hw.reshape(-1), vel, ek, e, s.reshape(-1).copy())
    assert np.abs(r["jtot"] - jt).max() <= 1e-10 * np.abs(jt).max()
    # write_check's invariants (heat_check.f90:13-49) from the same per-atom arrays
    tk, tp, te1, s9 = slot.ctx.heat_check(slot.ek.data_ptr())
    assert abs(tk - ek[:nr].sum()) <= 1e-12 * ek[:nr].sum()
    assert abs(tp - e.sum()) <= 1e-10 * scale
    e1 = re["e_sites"][0].sum()
    assert abs(te1 - e1) <= 1e-10 * abs(e1)
    sq = re["stress_qd_atoms"].reshape(3, 3)
    assert np.abs(s9 - sq).max() <= 1e-10 * np.abs(sq).max()
    assert np.abs(sq).max() > 0.0


def test_concurrent_slots_equal_alone(waterbox, engine):
    boxes = [waterbox.make_waterbox(64, 12.42, 4, seed=40 + i) for i in range(4)]
    slots = [engine.Slot(b) for b in boxes]
    alone = []
    for s in slots:
        r = s.evaluate()
        alone.append((r, s.forces()))
    print("4 slots concurrently, 3 evaluations each", flush=True)
    _, res = engine.run_concurrently(slots, 3)
    for s, (r0, f0), rs in zip(slots, alone, res):
        for r in rs:
            assert r["eng"] == r0["eng"] and r["vir"] == r0["vir"]
            assert np.array_equal(r["jtot"], r0["jtot"]) and np.array_equal(r["stress"], r0["stress"])
        assert np.array_equal(s.forces(), f0)


def test_overlapped_parts_equal_ordered(waterbox, engine):
    box = waterbox.make_waterbox(128, 15.65, 6, seed=7)
    a, b = engine.Slot(box, overlap=False), engine.Slot(box, overlap=True)
    print("ordered vs overlapped evaluation, 128 waters", flush=True)
    r0 = a.evaluate()
    f0 = a.forces()
    for _ in range(3):
        r = b.evaluate()
        for key in ("eng", "vir", "dipole_iterations"):
            assert r[key] == r0[key]
        for key in ("stress", "jkin", "jpot", "jtot"):
            assert np.array_equal(r[key], r0[key])
        assert np.array_equal(b.forces(), f0)
    b.close()


def test_2048_whole_path_size_independent(pkg, waterbox, engine):
    """The large-box workload at full size: 2048 waters on a 13^3 lattice (L = 39.43 A),
    Ewald kmax 24 (33 530 k vectors), rcut 9 A, every term and the heat flux, checked by
    what holds at any size: zero net force, per-atom energies summing to the energy (the
    long-range tail has no per-atom share), virial = -trace(stress) (the tail's -virlrc/3
    per diagonal is the caller's), heat flux linear in the velocities, the dipole solver
    converged, and a second evaluation bit-identical."""
    import torch
    box = waterbox.box_2048()
    slot = engine.Slot(box)
    print("2048 waters, whole path", flush=True)
    r = slot.evaluate()
    assert int(slot.ctx.elec_details()["k_vectors"]) == 33530
    f = slot.forces()
    assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).max()
    tk, tp, te1, _ = slot.ctx.heat_check(slot.ek.data_ptr())
    ew = r["eng_water"]
    assert abs(tp - (r["eng"] - ew[3])) <= 1e-10 * abs(r["eng_elec"])
    vw = r["vir_water"]
    tr = r["stress"][0] + r["stress"][4] + r["stress"][8]
    assert abs(tr + (r["vir"] - vw[3])) <= 1e-10 * abs(r["vir"])
    assert 0 < r["dipole_iterations"] < 200
    # J is linear in v for fixed per-atom kinetic energies
    v2 = 2.0 * slot.v
    torch.cuda.synchronize()
    jk, jp, jt = slot.ctx.heat_flux_dev(v2[0].data_ptr(), v2[1].data_ptr(), v2[2].data_ptr(), slot.ek.data_ptr())
    assert np.abs(np.array(jt) - 2.0 * r["jtot"]).max() <= 1e-13 * np.abs(r["jtot"]).max()
    r2 = slot.evaluate()
    assert r2["eng"] == r["eng"] and np.array_equal(r2["jtot"], r["jtot"]) and np.array_equal(r2["stress"], r["stress"])
    assert np.array_equal(slot.forces(), f)


FP32_BAR_REASON = (
    "FP32 polynomials miss the mixed-precision bar (RMS dF <= 1e-5 kcal/mol/A, |dJ|/|J| <= 1e-6) by about "
    "10^3: the terms of the 2B and 3B polynomials cancel by about 10^6 (sum |terms| / |E|), so FP32 rounding "
    "of the terms is amplified past the bar; re-centring the polynomials with the constant and linear terms "
    "in FP64 gains less than 3x (oracle/precision_study.py)")


@pytest.fixture(scope="module")
def fp32_vs_fp64(waterbox, engine):
    box = waterbox.box_512()
    a, b = engine.Slot(box, precision=0), engine.Slot(box, precision=3)
    ra, rb = a.evaluate(), b.evaluate()
    return ra, rb, a.forces(), b.forces()


@pytest.mark.xfail(strict=True, reason=FP32_BAR_REASON)
def test_fp32_polynomials_meet_the_mixed_precision_bar(fp32_vs_fp64):
    ra, rb, fa, fb = fp32_vs_fp64
    df = (fb - fa) / 418.4
    rms = float(np.sqrt((df ** 2).mean()))
    rel_dj = float(np.abs(rb["jtot"] - ra["jtot"]).max() / np.abs(ra["jtot"]).max())
    print(f"fp32 polynomials: rms dF {rms:.3e} kcal/mol/A, rel dJ {rel_dj:.2e}", flush=True)
    assert rms <= 1e-5 and rel_dj <= 1e-6


def test_fp32_polynomials_leave_the_fp64_parts(fp32_vs_fp64):
    """The FP32 mode changes only the polynomials: the electrostatics and the water terms,
    which do not depend on them, are bit-identical; the MB-pol energy moves."""
    ra, rb, fa, fb = fp32_vs_fp64
    assert rb["eng_elec"] == ra["eng_elec"]
    assert np.array_equal(rb["eng_water"], ra["eng_water"])
    assert rb["eng_mbpol"] != ra["eng_mbpol"]


@pytest.mark.parametrize("solver", [0, 1])
def test_no_read_of_unwritten_memory(pkg, waterbox, engine, solver, monkeypatch):
    """Every floating-point buffer filled with NaN at allocation (MBGPU_POISON=1): a kernel
    that read an element it did not write in the same evaluation would turn the results
    into NaN.  Results must equal those of a context on zero-filled memory, bit for bit
    (regression: the conjugate-gradient direction's M-site entries were read once)."""
    cases = [(waterbox.box_64(), pkg.elec_params.for_box(12.42), pkg.water_params.for_box(12.42))]
    c = waterbox.make_cluster(40, 3)
    cases.append((c, pkg.elec_params.for_cluster(c.L), pkg.water_params.for_cluster(c.L)))
    for box, e, w in cases:
        out = []
        for poison in ("0", "1"):
            monkeypatch.setenv("MBGPU_POISON", poison)
            s = engine.Slot(box, solver=solver, elec=e, water=w)
            r = s.evaluate()
            out.append((r, s.forces(), s.ctx.elec_details(), s.ctx.water_details(), s.ctx.decomposition()))
            del s
        (r0, f0, d0, w0, m0), (r1, f1, d1, w1, m1) = out
        assert r1["eng"] == r0["eng"] and np.array_equal(f1, f0) and np.array_equal(r1["jtot"], r0["jtot"])
        for k in ("atomic_energy", "heat_w", "s_ew1p", "chge", "vesp"):
            assert np.array_equal(d1[k], d0[k]), k
        for k in ("atomic_energy", "heat_w"):
            assert np.array_equal(w1[k], w0[k]), k
        for a, b in zip(m1, m0):
            assert np.array_equal(a, b)


def test_neighbour_rows_grow_and_repeat(pkg, waterbox, engine, monkeypatch):
    box = waterbox.box_64()
    monkeypatch.setenv("MBGPU_ELEC_ROWS", "64")     # sites within 6 A at liquid density: ~120
    small = engine.Slot(box)
    print("evaluation with 64-entry neighbour rows (grown to 128, then repeated)", flush=True)
    r1 = small.evaluate()
    monkeypatch.delenv("MBGPU_ELEC_ROWS")
    ref = engine.Slot(box)
    r0 = ref.evaluate()
    assert r1["eng"] == r0["eng"] and np.array_equal(r1["jtot"], r0["jtot"])
    assert np.array_equal(small.forces(), ref.forces())
    assert r1["dipole_iterations"] == r0["dipole_iterations"]


def test_total_forces_equal_the_host_sum_of_the_parts(pkg):
    # mbgpu_get_forces (engine.Slot.forces) against the sum of the three parts' details in
    # the same order, through both evaluation paths
    engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
    box = pkg.waterbox.box_64()
    for fused in (True, False):
        slot = engine.Slot(box, fused=fused)
        for _ in range(3):
            slot.evaluate()
        nr = box.nreal
        _, _, fm = slot.ctx.decomposition()
        fe = slot.ctx.elec_details()["f_sites"]
        fw = slot.ctx.water_details()["f"]
        assert np.array_equal(slot.forces(), fm[:nr] + fe[:nr] + fw)
    with pytest.raises(pkg.MbgpuError, match="no complete evaluation"):
        pkg.Context(8, 32).total_forces()


def test_dedicated_queues_give_the_same_results(pkg):
    # mbgpu_create_queues(..., 1): the same kernels on streams with queues of their own
    engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
    box = pkg.waterbox.box_64()
    a = engine.Slot(box)
    b = engine.Slot(box, dedicated_queues=True)
    for _ in range(3):
        ra, rb = a.evaluate(), b.evaluate()
        for k in ("eng", "vir", "dipole_iterations"):
            assert ra[k] == rb[k], k
        assert np.array_equal(ra["jtot"], rb["jtot"]) and np.array_equal(ra["stress"], rb["stress"])
        assert np.array_equal(a.forces(), b.forces())
    lib = pkg.mbgpu.load_library()
    assert not lib.mbgpu_create_queues(8, 32, 2)
    assert "dedicated is 0 or 1" in lib.mbgpu_last_error().decode()
