"""GPU parity of the MB-pol force path through the C ABI (libmbgpu.so) against the oracle.

Bars (FP64 path):
  * dimer and trimer sets bit-exact against the oracle's predicate sets;
  * energies, virial, stress: relative 1e-10;
  * forces, per-atom energies, per-atom heat-flux tensors: max |diff| <= 1e-10 * max |value|;
  * heat flux: relative 1e-10 per component against the oracle (dense force matrix at 64
    waters, per-atom tensors at 512);
  * leaves: within 1e-15 * S of the reference's golden vectors, S = sum |a_b monomial|
    (the cancellation scale of the polynomial; see test_oracle_leaves.py).
At 2048 waters (beyond the oracle's few-second range) the checks are size-independent:
the complete dimer and trimer sets against the predicates evaluated for all pairs (the GPU
scanning every water and, in a second run, searching six link cells per axis), zero net
force, energy decomposition sums,
virial = -trace(stress), determinism.
"""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def gpu_forces(pkg, box, ctx=None):
    ctx = ctx or pkg.Context(box.nw, box.natms)
    f = np.zeros((3, box.natms))
    stress = np.zeros(9)
    e, v = ctx.mbpol_forces(box.imcon, box.cell, 3 * box.nw, box.listttm2, box.x, box.y, box.z, f[0], f[1], f[2],
                            0.0, 0.0, stress)
    ae, hw, fonly = ctx.decomposition()
    d, t = ctx.lists()
    return ctx, dict(f=f.T.copy(), eng=e, vir=v, stress=stress, atomic_energy=ae, heat_w=hw, dimers=d, trimers=t)


def test_cutoffs(pkg):
    assert pkg.mbpol_2b_cutoff() == 6.5 and pkg.mbpol_3b_cutoff() == 4.5


def test_leaf_2b_golden(pkg, golden):
    for w, E, g, S in zip(golden["d_w"], golden["d_E"], golden["d_g"], golden["d_S"]):
        g1, g2 = np.full(9, 3.0), np.full(9, 3.0)
        e, g1, g2 = pkg.mbpol_2b_poly(w[:9], w[9:], g1, g2)
        if np.linalg.norm(w[:3] - w[9:12]) > 6.5:
            assert e == 0.0 and np.all(g1 == 3.0) and np.all(g2 == 3.0)
            continue
        assert abs(e - E) <= 1e-15 * S
        assert np.abs(np.concatenate([g1, g2]) - g).max() <= 1e-15 * S


def test_leaf_3b_golden_accumulates(pkg, golden):
    for w, E, g, S in zip(golden["t_w"], golden["t_E"], golden["t_g"], golden["t_S"]):
        g0 = [np.full(9, 0.5) for _ in range(3)]
        e, g1, g2, g3 = pkg.mbpol_3b_poly(w[:9], w[9:18], w[18:], *g0)
        assert abs(e - E) <= 1e-15 * S
        assert np.abs(np.concatenate([g1, g2, g3]) - 0.5 - g).max() <= 1e-15 * S


@pytest.fixture(scope="module")
def box64(waterbox):
    return waterbox.box_64()


@pytest.fixture(scope="module")
def box512(waterbox):
    return waterbox.box_512()


@pytest.fixture(scope="module")
def ora64(box64):
    return o.mbpol_forces(box64, compat_inplace=1, dense=True)


@pytest.fixture(scope="module")
def ora512(box512):
    return o.mbpol_forces(box512, compat_inplace=1)


def check_against_oracle(g, r):
    assert sorted(map(tuple, g["dimers"].tolist())) == sorted(map(tuple, r["dimers"].tolist()))
    gt = sorted(map(tuple, g["trimers"].tolist()))
    rt = sorted(map(tuple, r["trimers"].tolist()))
    assert gt == rt
    assert abs(g["eng"] - r["eng"]) <= 1e-10 * abs(r["eng"])
    assert abs(g["vir"] - r["vir"]) <= 1e-10 * abs(r["vir"])
    assert rel(g["stress"], r["stress"]) <= 1e-10
    assert rel(g["f"], r["f"]) <= 1e-10
    assert rel(g["atomic_energy"], r["atomic_energy"]) <= 1e-10
    assert rel(g["heat_w"], r["heat_w"]) <= 1e-10


def test_forces_64_match_oracle(pkg, box64, ora64):
    _, g = gpu_forces(pkg, box64)
    check_against_oracle(g, ora64)


def test_forces_512_match_oracle(pkg, box512, ora512):
    _, g = gpu_forces(pkg, box512)
    check_against_oracle(g, ora512)


def test_heat_flux_64_matches_dense_oracle(pkg, box64, ora64):
    ctx, g = gpu_forces(pkg, box64)
    vel = box64.vel
    ek = 0.5 * box64.mass * (vel ** 2).sum(1)
    jk, jp, jt = ctx.heat_flux(vel[:, 0], vel[:, 1], vel[:, 2], ek)
    rk, rp, rt = o.heat_flux_dense(box64, ora64["force_matrix"], vel, ek, ora64["atomic_energy"])
    assert rel(jk, rk) <= 1e-10 and rel(jp, rp) <= 1e-10 and rel(jt, rt) <= 1e-10


def test_heat_flux_512_with_extra_terms(pkg, box512, ora512):
    ctx, g = gpu_forces(pkg, box512)
    rng = np.random.default_rng(11)
    vel = box512.vel
    ek = 0.5 * box512.mass * (vel ** 2).sum(1)
    ep2 = rng.normal(size=box512.natms)
    w2 = rng.normal(size=(box512.natms, 9))
    sew = rng.normal(size=(box512.natms, 9))
    jk, jp, jt = ctx.heat_flux(vel[:, 0], vel[:, 1], vel[:, 2], ek, ep2, w2, sew)
    rk, rp, rt = o.heat_flux_tensor(box512, ora512["heat_w"].reshape(-1) + w2.reshape(-1), vel, ek,
                                    ora512["atomic_energy"] + ep2, sew.reshape(-1).copy())
    assert rel(jk, rk) <= 1e-10 and rel(jp, rp) <= 1e-10 and rel(jt, rt) <= 1e-10


def test_accumulates_into_caller_arrays(pkg, box64, ora64):
    ctx = pkg.Context(box64.nw, box64.natms)
    f = np.ones((3, box64.natms))
    stress = np.full(9, 2.0)
    e, v = ctx.mbpol_forces(1, box64.cell, 3 * box64.nw, box64.listttm2, box64.x, box64.y, box64.z, f[0], f[1], f[2],
                            5.0, 7.0, stress)
    assert abs(e - 5.0 - ora64["eng"]) <= 1e-10 * abs(ora64["eng"])
    assert abs(v - 7.0 - ora64["vir"]) <= 1e-10 * abs(ora64["vir"])
    assert rel(f.T - 1.0, ora64["f"]) <= 1e-10
    assert rel(stress - 2.0, ora64["stress"]) <= 1e-10


def test_deterministic_and_reusable(pkg, box512):
    ctx, a = gpu_forces(pkg, box512)
    _, b = gpu_forces(pkg, box512, ctx)
    for k in ("f", "atomic_energy", "heat_w", "stress"):
        assert np.array_equal(a[k], b[k]), k
    assert a["eng"] == b["eng"] and a["vir"] == b["vir"]


def test_unsupported_imcon_is_an_error(pkg, box64):
    ctx = pkg.Context(box64.nw, box64.natms)
    f = np.zeros((3, box64.natms))
    with pytest.raises(pkg.MbgpuError, match="imcon"):
        ctx.mbpol_forces(4, box64.cell, 3 * box64.nw, box64.listttm2, box64.x, box64.y, box64.z, f[0], f[1], f[2])


def test_bad_list_is_an_error(pkg, box64):
    ctx = pkg.Context(box64.nw, box64.natms)
    f = np.zeros((3, box64.natms))
    lst = box64.listttm2.copy()
    lst[5] = box64.natms + 10
    with pytest.raises(pkg.MbgpuError, match="listttm2"):
        ctx.mbpol_forces(1, box64.cell, 3 * box64.nw, lst, box64.x, box64.y, box64.z, f[0], f[1], f[2])


def test_bad_device_list_is_an_error(pkg, box64):
    """A device listttm2 is checked on the device before an atom array is indexed with it
    (the list of zeros is what a late clear of a serving slot's list left behind); the
    context evaluates a valid list unchanged afterwards."""
    import torch
    dev = torch.device("cuda:0")
    T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    x, y, z = T(box64.x), T(box64.y), T(box64.z)
    ctx = pkg.Context(box64.nw, box64.natms)
    good = T(box64.listttm2, torch.int32)
    torch.cuda.synchronize()
    e0, v0, s0 = ctx.mbpol_forces_dev(box64.imcon, box64.cell, good.data_ptr(), x.data_ptr(), y.data_ptr(),
                                      z.data_ptr())
    swapped = box64.listttm2.copy()
    swapped[[4, 5]] = swapped[[5, 4]]
    beyond = box64.listttm2.copy()
    beyond[-3:] = [box64.natms - 1, box64.natms, box64.natms + 1]
    for name, bad in (("zeros", np.zeros_like(box64.listttm2)), ("swapped", swapped), ("beyond", beyond)):
        print("bad device list:", name, flush=True)
        lst = T(bad, torch.int32)
        torch.cuda.synchronize()
        with pytest.raises(pkg.MbgpuError, match="listttm2"):
            ctx.mbpol_forces_dev(box64.imcon, box64.cell, lst.data_ptr(), x.data_ptr(), y.data_ptr(), z.data_ptr())
    e1, v1, s1 = ctx.mbpol_forces_dev(box64.imcon, box64.cell, good.data_ptr(), x.data_ptr(), y.data_ptr(),
                                      z.data_ptr())
    assert e1 == e0 and v1 == v0 and np.array_equal(s1, s0)


@pytest.mark.parametrize("cells", [False, True])
def test_2048_size_independent_properties(pkg, waterbox, cells, monkeypatch):
    box = waterbox.box_2048()
    if cells:
        monkeypatch.setenv("MBGPU_CELLS_MIN", "1024")   # the O-O lists through link cells
    ctx, g = gpu_forces(pkg, box)
    d, t = g["dimers"], g["trimers"]
    assert len(d) > 30000 and len(t) > 100000
    assert np.all(d[:, 0] < d[:, 1])
    assert np.all((t[:, 0] < t[:, 1]) & (t[:, 1] < t[:, 2]))
    assert len(set(map(tuple, d.tolist()))) == len(d)
    assert len(set(map(tuple, t.tolist()))) == len(t)
    f = g["f"]
    assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).max()
    assert abs(g["atomic_energy"].sum() - g["eng"]) <= 1e-10 * abs(g["eng"])
    s = g["stress"]
    assert abs(s[0] + s[4] + s[8] + g["vir"]) <= 1e-10 * abs(g["vir"])
    # the complete sets against the predicates of mbpol.f90:496-503 (dimers) and 843-869
    # (trimers: at least two O-O edges <= r3f), evaluated for all pairs in the same FP64
    # order (lower index minus higher, d - L nint(d / L), (x^2 + y^2) + z^2)
    O = box.xyz[:box.nreal:3]
    dd = O[:, None, :] - O[None, :, :]
    dd = dd - box.L * np.round(dd * (1.0 / box.L))
    r2 = (dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1]) + dd[..., 2] * dd[..., 2]
    iu, ju = np.triu_indices(box.nw, 1)
    r2u = r2[iu, ju]
    pair = ~(r2u > 6.5 * 6.5)
    assert set(map(tuple, d.tolist())) == set(zip(iu[pair].tolist(), ju[pair].tolist()))
    short = ~(np.sqrt(r2u) > 4.5)
    nb = [[] for _ in range(box.nw)]
    for a_, b_ in zip(iu[short].tolist(), ju[short].tolist()):
        nb[a_].append(b_)
        nb[b_].append(a_)
    tri = set()
    for c_ in range(box.nw):
        ns = sorted(nb[c_])
        for u in range(len(ns)):
            for v in range(u + 1, len(ns)):
                tri.add(tuple(sorted((c_, ns[u], ns[v]))))
    assert set(map(tuple, t.tolist())) == tri
    _, g2 = gpu_forces(pkg, box, ctx)
    assert np.array_equal(g["f"], g2["f"])
