"""Complete force evaluations through engine.Slot, alone and concurrently.

* One slot on the 64-water box: total energy, virial, forces and heat flux equal the sum of
  the oracle's MB-pol, electrostatics and water terms within 1e-10.
* Several slots driven concurrently from host threads (their kernels overlap on the GPU):
  every result is bit-identical to the same slot evaluated alone.
* The electrostatics on the context's second stream from a helper thread, beside MB-pol and
  the water terms (Slot overlap), give bit-identical results to the three parts in order.
* The FP32-polynomial mode stays within its measured error envelope against FP64.
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


def test_fp32_polynomials_error_envelope(waterbox, engine):
    """FP32 polynomials: RMS force error 1e-2 kcal/mol/A class at 512 waters (the polynomial
    conditioning makes 1e-5 unreachable in FP32); the FP64 parts are untouched."""
    box = waterbox.box_512()
    a, b = engine.Slot(box, precision=0), engine.Slot(box, precision=3)
    ra, rb = a.evaluate(), b.evaluate()
    df = (b.forces() - a.forces()) / 418.4
    rms = float(np.sqrt((df ** 2).mean()))
    print(f"fp32 polynomials: rms dF {rms:.3e} kcal/mol/A", flush=True)
    assert 0.0 < rms < 0.05
    assert abs(rb["eng_elec"] - ra["eng_elec"]) <= 1e-6 * abs(ra["eng_elec"])
    assert np.array_equal(rb["eng_water"], ra["eng_water"])


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
